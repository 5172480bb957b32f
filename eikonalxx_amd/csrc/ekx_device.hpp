// ekx_device.hpp -- descriptors shared by the host launch code and the HIP
// kernels of the fast-sweeping solver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ekx
{

// Sentinel of an unreached node: HUGE in include/private/solver3d.hpp:10 and
// include/private/solver2d.hpp:16.
constexpr double kHuge = 1.e10;

// Tile footprint of the column kernel: one 64-lane wave owns an 8 x 8 block
// of (x, y) nodes and marches through z.
constexpr int kColumnTile = 8;
constexpr int kColumnHalo = kColumnTile + 1;          // 9: tile + upwind halo
constexpr int kColumnPlane = kColumnHalo*kColumnHalo; // 81 values per plane
constexpr int kColumnSkew = 2*(kColumnTile - 1);      // 14: lane a+b spread
// LDS ring of planes per wave.  A plane is stored at step q, read by the
// lanes through step q + 14 and written back at step q + 15, so 16 slots
// are enough.
constexpr int kColumnRing = 16;

// Grid constants of the non-uniform local solver
// (include/private/solver3d.hpp:247-261).
struct NonUniform3D
{
    double dxInv, dyInv, dzInv;
    double dx2Inv, dy2Inv, dz2Inv;
    double dx_dz, dz_dx, dy_dz, dz_dy, dx_dy, dy_dx;
    double dx2_p_dz2_inv, dy2_p_dz2_inv, dx2_p_dy2_inv;
};

// One independent 3D problem (geometry + model + source) of a batch.
struct Problem3D
{
    int nx, ny, nz;
    int radius;                 // factored-eikonal radius (grid points)
    int srcCellX, srcCellY, srcCellZ;
    int uniform;                // dx == dy == dz within 100 eps
    double dx, dy, dz;
    double srcOffX, srcOffY, srcOffZ;  // source - origin (m)
    double srcSlowness;         // slowness of the source cell
    int64_t tOffset;            // first node of this problem's field
    int64_t sOffset;            // first cell of this problem's model
    // Scaled-integer storage (uint32 codes): travel time = code * tStep,
    // slowness = code * sStep; tScale = 1 / tStep.  Unused otherwise.
    double tStep, tScale, sStep;
    NonUniform3D nu;
};

// A column chunk (8 x 8 footprint, one z-chunk) of one problem.
struct ColumnItem
{
    int problem;
    int bx, by, bz;             // column and chunk index, permuted frame
};

// A receiver located in one problem's grid (Station3D, src/station3d.cpp:
// 158-291): offset from the model origin and the cell that holds it.
struct StationLoc3D
{
    double off[3];
    int cell[3];
    int pad;
};

// One problem of a 2D batch.
struct Problem2D
{
    int nx, nz;
    int radius;
    int srcCellX, srcCellZ;
    int uniform;
    int nSweeps;
    int pad;
    double dx, dz;
    double srcOffX, srcOffZ;
    double srcSlowness;
    double dx_dz, dz_dx, cosTheta, sinTheta;  // solver2d.hpp:41-46
    int64_t tOffset;
    int64_t sOffset;
};

} // namespace ekx
