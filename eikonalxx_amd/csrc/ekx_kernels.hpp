// ekx_kernels.hpp -- host-side launchers of the HIP kernels.
#pragma once
#include "ekx_device.hpp"

namespace ekx
{

// Waves per workgroup of the column kernel: LDS per wave is
// 16 slots x 2 x 84 values, so FP32 blocks of 4 waves and FP64 blocks of 2
// waves both use 43 KB.
template <typename TS> constexpr int column_waves()
{
    return sizeof(TS) == 4 ? 4 : 2;
}

template <typename TS>
hipError_t launch_init3d(const Problem3D *dProblems, int nProblems,
                         int64_t maxNodes, TS *T, hipStream_t stream);
template <typename TS>
hipError_t launch_hyperplane3d(const Problem3D *dProblems, int problem,
                               int sweep, int initialize, int level, int kLo,
                               int nPairs, TS *T, const TS *S,
                               hipStream_t stream);
template <typename TS>
hipError_t launch_columns3d(const Problem3D *dProblems,
                            const ColumnItem *dItems, int nItems, int sweep,
                            int initialize, TS *T, const TS *S,
                            hipStream_t stream);
hipError_t launch_nodal_to_cell3d(int nx, int ny, int nz, const double *v,
                                  int ordering, double *slow,
                                  hipStream_t stream);
hipError_t launch_cell_velocity_to_slowness3d(int nx, int ny, int nz,
                                              const double *v, int ordering,
                                              double *slow,
                                              hipStream_t stream);
hipError_t launch_convert(int64_t n, const double *in, float *out,
                          hipStream_t stream);
hipError_t launch_convert(int64_t n, const float *in, double *out,
                          hipStream_t stream);

hipError_t launch_solve2d(const Problem2D *dProblems, int nProblems,
                          double *T, const double *S, hipStream_t stream);
hipError_t launch_stations2d(const Problem2D *dProblem, int n,
                             const double *xOff, const double *zOff,
                             const int *cellX, const int *cellZ,
                             const double *T, double *out,
                             hipStream_t stream);
hipError_t launch_nodal_to_cell2d(int nx, int nz, const double *v,
                                  int ordering, double *slow,
                                  hipStream_t stream);

} // namespace ekx
