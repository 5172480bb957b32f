// ekx_kernels3d.hip -- HIP kernels of the 3D fast-sweeping solver (gfx950).
//
// Sweep k (0..7) visits nodes in the order of Solver3DSweep<k>::updateFSM
// (include/private/solver3d.hpp:150-326).  In the sweep's permuted frame
// (solverUtilities3d.hpp:1197-1257) every sweep runs towards +x, +y, +z and
// a node reads only its 7 upwind neighbours, which lie on earlier
// hyperplanes jx+jy+jz.  Any schedule that finishes those neighbours first
// therefore reproduces the serial loop exactly; both kernels below are such
// schedules.
#include "ekx_kernels.hpp"
#include "ekx_local3d.hpp"

namespace ekx
{

// First node of each permuted axis that a sweep updates: 1 for refinement
// sweeps (getLoopLimits, solverUtilities3d.hpp:944-1052), the source octant
// for initialization sweeps (:1075-1184).
struct Octant
{
    int jx0, jy0, jz0;
};

__device__ __forceinline__ Octant sweep_octant(const Problem3D &p, int sweep,
                                               int initialize)
{
    Octant o{1, 1, 1};
    if (initialize)
    {
        int negX = sweep & 1, negY = (sweep >> 1) & 1, negZ = (sweep >> 2) & 1;
        o.jx0 = negX ? max(p.nx - 2 - p.srcCellX, 1) : max(1, p.srcCellX);
        o.jy0 = negY ? max(p.ny - 2 - p.srcCellY, 1) : max(1, p.srcCellY);
        o.jz0 = negZ ? max(p.nz - 2 - p.srcCellZ, 1) : max(1, p.srcCellZ);
    }
    return o;
}

// ---------------------------------------------------------------------------
// Source initialisation (Solver3DImpl::initializeTravelTimes,
// src/solver3d.cpp:21-84): HUGE everywhere, the analytic time at the 8
// corners of the source cell.  blockIdx.y is the problem.
// ---------------------------------------------------------------------------
template <typename TS>
__global__ void __launch_bounds__(256)
init_travel_times3d(const Problem3D *__restrict__ problems,
                    TS *__restrict__ T)
{
    const Problem3D &p = problems[blockIdx.y];
    const int64_t nxy = (int64_t) p.nx*p.ny;
    const int64_t n = nxy*p.nz;
    TS *Tp = T + p.tOffset;
    for (int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x*blockDim.x)
    {
        int iz = (int) (i/nxy);
        int rem = (int) (i - (int64_t) iz*nxy);
        int iy = rem/p.nx;
        int ix = rem - iy*p.nx;
        int ox = ix - p.srcCellX, oy = iy - p.srcCellY, oz = iz - p.srcCellZ;
        double v = kHuge;
        if (ox >= 0 && ox <= 1 && oy >= 0 && oy <= 1 && oz >= 0 && oz <= 1)
        {
            v = analytic_time3d(ix, iy, iz, p.dx, p.dy, p.dz, p.srcOffX,
                                p.srcOffY, p.srcOffZ, p.srcSlowness);
        }
        Tp[i] = (TS) v;
    }
}

// ---------------------------------------------------------------------------
// Hyperplane kernel: all nodes of one level L = jx+jy+jz of one sweep (the
// level-set method of src/graph3d.cpp:119-221, with the node set computed
// in closed form instead of stored lists).  Threads enumerate (jz, jy)
// pairs; jx follows from the level.  Reads and writes global memory
// directly.
// ---------------------------------------------------------------------------
template <typename TS>
__global__ void __launch_bounds__(256)
sweep_hyperplane3d(const Problem3D *__restrict__ problems, int problem,
                   int sweep, int initialize, int level, int kLo, int nPairs,
                   TS *__restrict__ T, const TS *__restrict__ S)
{
    const int pid = blockIdx.x*blockDim.x + threadIdx.x;
    if (pid >= nPairs) { return; }
    const Problem3D &p = problems[problem];
    const Octant o = sweep_octant(p, sweep, initialize);
    const int nyRange = p.ny - o.jy0;
    const int jz = kLo + pid/nyRange;
    const int jy = o.jy0 + pid%nyRange;
    const int jx = level - jy - jz;
    if (jx < o.jx0 || jx > p.nx - 1) { return; }
    const int negX = sweep & 1, negY = (sweep >> 1) & 1, negZ = (sweep >> 2) & 1;
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int ncx = nx - 1, ncy = ny - 1, ncz = nz - 1;
    // natural indices of the node and of its upwind neighbour per axis
    const int ix = negX ? nx - 1 - jx : jx;
    const int iy = negY ? ny - 1 - jy : jy;
    const int iz = negZ ? nz - 1 - jz : jz;
    const int ixU = negX ? ix + 1 : ix - 1;
    const int iyU = negY ? iy + 1 : iy - 1;
    const int izU = negZ ? iz + 1 : iz - 1;
    // cells: permuted index j-1 (home) and min(j, nCell-1) (offset)
    auto cellx = [&](int pc) { return negX ? ncx - 1 - pc : pc; };
    auto celly = [&](int pc) { return negY ? ncy - 1 - pc : pc; };
    auto cellz = [&](int pc) { return negZ ? ncz - 1 - pc : pc; };
    const int cxH = cellx(jx - 1), cxO = cellx(min(jx, ncx - 1));
    const int cyH = celly(jy - 1), cyO = celly(min(jy, ncy - 1));
    const int czH = cellz(jz - 1), czO = cellz(min(jz, ncz - 1));
    const TS *Sp = S + p.sOffset;
    auto sAt = [&](int cx, int cy, int cz)
    {
        return Sp[((int64_t) cz*ncy + cy)*ncx + cx];
    };
    TS *Tp = T + p.tOffset;
    auto tIdx = [&](int x, int y, int z)
    {
        return ((int64_t) z*ny + y)*nx + x;
    };
    const TS s0 = sAt(cxH, cyH, czH);
    const TS s1 = sAt(cxO, cyH, czH);
    const TS s2 = sAt(cxO, cyO, czH);
    const TS s3 = sAt(cxH, cyO, czH);
    const TS s4 = sAt(cxH, cyH, czO);
    const TS s5 = sAt(cxO, cyH, czO);
    const TS s7 = sAt(cxH, cyO, czO);
    const int64_t i0 = tIdx(ix, iy, iz);
    const TS t0 = Tp[i0];
    const TS t1 = Tp[tIdx(ixU, iy, iz)];
    const TS t2 = Tp[tIdx(ixU, iyU, iz)];
    const TS t3 = Tp[tIdx(ix, iyU, iz)];
    const TS t4 = Tp[tIdx(ix, iy, izU)];
    const TS t5 = Tp[tIdx(ixU, iy, izU)];
    const TS t6 = Tp[tIdx(ixU, iyU, izU)];
    const TS t7 = Tp[tIdx(ix, iyU, izU)];
    const Sweep3D sw = make_sweep3d(sweep);
    const TS tUpd = local_update3d<TS>(p, sw, ix, iy, iz, s0, s1, s2, s3, s4,
                                       s5, s7, t1, t2, t3, t4, t5, t6, t7);
    if (tUpd < t0) { Tp[i0] = tUpd; }
}

// ---------------------------------------------------------------------------
// Column kernel.  One 64-lane wave owns an 8 x 8 (jx, jy) column of one
// problem and marches through z.  Lane (a, b) works on plane
// q = step - (a + b) (relative to the first plane of the sweep), so at every
// step the lanes hold nodes on one hyperplane and a lane's x- and
// y-neighbours on the same plane were finished one or two steps earlier by
// the neighbouring lanes.  Planes are staged in a 16-slot LDS ring of
// 9 x 9 values (the 8 x 8 tile plus the upwind halo row and column, which
// the previous column diagonal has already finished); the z-neighbours stay
// in registers.  Planes are prefetched DL steps ahead and written back,
// row-coalesced, once the last lane has finished them.  The waves of a
// block are independent: there is no block barrier.
//
// Launch: one launch per column diagonal bx + by = d.  Columns on one
// diagonal never read or write each other's nodes; their halos belong to
// the diagonals before.
// ---------------------------------------------------------------------------
template <typename TS, int WAVES, int DL>
__global__ void __launch_bounds__(64*WAVES)
sweep_columns3d(const Problem3D *__restrict__ problems,
                const ColumnItem *__restrict__ items, int nItems, int sweep,
                int initialize, TS *__restrict__ T,
                const TS *__restrict__ S)
{
    constexpr int kSlot = kColumnPlane + 3;   // 84: pad the 81 values
    __shared__ TS ring[WAVES][kColumnRing][2][kSlot];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int itemIndex = blockIdx.x*WAVES + wave;
    if (itemIndex >= nItems) { return; }
    const ColumnItem item = items[itemIndex];
    const int pIndex = __builtin_amdgcn_readfirstlane(item.problem);
    const Problem3D &p = problems[pIndex];
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int ncx = nx - 1, ncy = ny - 1, ncz = nz - 1;
    const int negX = sweep & 1, negY = (sweep >> 1) & 1, negZ = (sweep >> 2) & 1;
    const Octant o = sweep_octant(p, sweep, initialize);
    const int X0 = __builtin_amdgcn_readfirstlane(item.bx)*kColumnTile;
    const int Y0 = __builtin_amdgcn_readfirstlane(item.by)*kColumnTile;
    if (X0 + kColumnTile - 1 < o.jx0 || Y0 + kColumnTile - 1 < o.jy0)
    {
        return;   // column entirely outside the source octant
    }
    const int P0 = o.jz0 - 1;             // first plane: read only
    const int nPlanes = nz - P0;
    const int a = lane & 7, b = lane >> 3;
    const int jx = X0 + a, jy = Y0 + b;
    const bool valid = jx < nx && jy < ny;
    const bool upd = valid && jx >= o.jx0 && jy >= o.jy0;
    const int ix = negX ? nx - 1 - jx : jx;
    const int iy = negY ? ny - 1 - jy : jy;
    const Sweep3D sw = make_sweep3d(sweep);
    TS *Tp = T + p.tOffset;
    const TS *Sp = S + p.sOffset;
    const int64_t nxy = (int64_t) nx*ny;
    const int64_t ncxy = (int64_t) ncx*ncy;

    // Ring elements this lane loads: e0 = lane, e1 = lane + 64 (< 81).
    const bool has1 = lane + 64 < kColumnPlane;
    int tXY0 = -1, tXY1 = -1, sXY0, sXY1;
    {
        auto setup = [&](int e, int &tXY, int &sXY)
        {
            int c = e%kColumnHalo, r = e/kColumnHalo;
            int gx = X0 - 1 + c, gy = Y0 - 1 + r;
            if (gx >= 0 && gx < nx && gy >= 0 && gy < ny)
            {
                int ax = negX ? nx - 1 - gx : gx;
                int ay = negY ? ny - 1 - gy : gy;
                tXY = ay*nx + ax;
            }
            int pcx = min(max(gx, 0), ncx - 1);
            int pcy = min(max(gy, 0), ncy - 1);
            int cx = negX ? ncx - 1 - pcx : pcx;
            int cy = negY ? ncy - 1 - pcy : pcy;
            sXY = cy*ncx + cx;
        };
        setup(lane, tXY0, sXY0);
        sXY1 = sXY0;
        if (has1) { setup(lane + 64, tXY1, sXY1); }
    }
    const TS huge = (TS) kHuge;
    auto loadPlane = [&](int q, TS &vT0, TS &vT1, TS &vS0, TS &vS1)
    {
        const int jz = P0 + q;
        const int az = negZ ? nz - 1 - jz : jz;
        const int pcz = min(jz, ncz - 1);
        const int cz = negZ ? ncz - 1 - pcz : pcz;
        const TS *Tz = Tp + (int64_t) az*nxy;
        const TS *Sz = Sp + (int64_t) cz*ncxy;
        vT0 = tXY0 >= 0 ? Tz[tXY0] : huge;
        vS0 = Sz[sXY0];
        vT1 = huge;
        vS1 = vS0;
        if (has1)
        {
            vT1 = tXY1 >= 0 ? Tz[tXY1] : huge;
            vS1 = Sz[sXY1];
        }
    };

    TS bT0[DL], bT1[DL], bS0[DL], bS1[DL];
#pragma unroll
    for (int u = 0; u < DL; ++u)
    {
        if (u < nPlanes) { loadPlane(u, bT0[u], bT1[u], bS0[u], bS1[u]); }
    }
    TS pt4 = huge, pt5 = huge, pt6 = huge, pt7 = huge;
    TS ps00 = 0, ps10 = 0, ps01 = 0, ps11 = 0;
    const int own = (b + 1)*kColumnHalo + a + 1;
    const int nSteps = nPlanes + kColumnSkew + 1;
    TS(*myRing)[2][kSlot] = ring[wave];
    for (int t0 = 0; t0 < nSteps; t0 += DL)
    {
#pragma unroll
        for (int u = 0; u < DL; ++u)
        {
            const int tc = t0 + u;
            if (tc >= nSteps) { break; }
            // (a) the plane prefetched DL steps ago enters the ring
            if (tc < nPlanes)
            {
                TS *dT = myRing[tc & (kColumnRing - 1)][0];
                TS *dS = myRing[tc & (kColumnRing - 1)][1];
                dT[lane] = bT0[u];
                dS[lane] = bS0[u];
                if (has1)
                {
                    dT[lane + 64] = bT1[u];
                    dS[lane + 64] = bS1[u];
                }
            }
            // (b) prefetch plane tc + DL into the freed registers
            if (tc + DL < nPlanes)
            {
                loadPlane(tc + DL, bT0[u], bT1[u], bS0[u], bS1[u]);
            }
            // (c) every lane advances its node on plane tc - (a + b)
            const int q = tc - (a + b);
            if (valid && q >= 0 && q < nPlanes)
            {
                TS *sT = myRing[q & (kColumnRing - 1)][0];
                const TS *sS = myRing[q & (kColumnRing - 1)][1];
                TS tOwn = sT[own];
                const TS t1 = sT[own - 1];
                const TS t2 = sT[own - 1 - kColumnHalo];
                const TS t3 = sT[own - kColumnHalo];
                const TS c00 = sS[b*kColumnHalo + a];
                const TS c10 = sS[b*kColumnHalo + a + 1];
                const TS c01 = sS[(b + 1)*kColumnHalo + a];
                const TS c11 = sS[(b + 1)*kColumnHalo + a + 1];
                if (q > 0 && upd)
                {
                    const int jz = P0 + q;
                    const int iz = negZ ? nz - 1 - jz : jz;
                    const TS tUpd = local_update3d<TS>(
                        p, sw, ix, iy, iz, ps00, ps10, ps11, ps01, c00, c10,
                        c01, t1, t2, t3, pt4, pt5, pt6, pt7);
                    if (tUpd < tOwn)
                    {
                        tOwn = tUpd;
                        sT[own] = tUpd;
                    }
                }
                pt4 = tOwn;
                pt5 = t1;
                pt6 = t2;
                pt7 = t3;
                ps00 = c00;
                ps10 = c10;
                ps01 = c01;
                ps11 = c11;
            }
            // (d) the plane the last lane finished a step ago goes home
            const int w = tc - (kColumnSkew + 1);
            if (w >= 1 && w < nPlanes && upd)
            {
                const int jz = P0 + w;
                const int az = negZ ? nz - 1 - jz : jz;
                Tp[(int64_t) az*nxy + (int64_t) iy*nx + ix]
                    = myRing[w & (kColumnRing - 1)][0][own];
            }
            // LDS traffic of this step must not be reordered with the next
            // step's: other lanes of the wave consume it.
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// ---------------------------------------------------------------------------
// Model preparation (src/model3d.cpp:45-151 interpolate3d, :316-375).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
nodal_to_cell3d(int nx, int ny, int nz, const double *__restrict__ v,
                int ordering, double *__restrict__ slow)
{
    const int64_t ncx = nx - 1, ncy = ny - 1, ncz = nz - 1;
    const int64_t nCell = ncx*ncy*ncz;
    for (int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x; i < nCell;
         i += (int64_t) gridDim.x*blockDim.x)
    {
        int64_t iz = i/(ncx*ncy);
        int64_t rem = i - iz*ncx*ncy;
        int64_t iy = rem/ncx;
        int64_t ix = rem - iy*ncx;
        int64_t idx1, idx2, idx3, idx4, idx5, idx6, idx7, idx8;
        if (ordering == 0)
        {
            idx1 = iz*(ncx + 1)*(ncy + 1) + iy*(ncx + 1) + ix;
            idx2 = idx1 + 1;
            idx3 = idx1 + (ncx + 1);
            idx4 = idx3 + 1;
            idx5 = idx1 + (ncx + 1)*(ncy + 1);
            idx6 = idx2 + (ncx + 1)*(ncy + 1);
            idx7 = idx3 + (ncx + 1)*(ncy + 1);
            idx8 = idx4 + (ncx + 1)*(ncy + 1);
        }
        else
        {
            idx1 = ix*(ncz + 1)*(ncy + 1) + iy*(ncz + 1) + iz;
            idx2 = idx1 + 1;
            idx3 = idx1 + (ncz + 1);
            idx4 = idx3 + 1;
            idx5 = idx1 + (ncz + 1)*(ncy + 1);
            idx6 = idx2 + (ncz + 1)*(ncy + 1);
            idx7 = idx3 + (ncz + 1)*(ncy + 1);
            idx8 = idx4 + (ncz + 1)*(ncy + 1);
        }
        double sum1 = v[idx1] + v[idx2];
        double sum2 = v[idx3] + v[idx4];
        double sum3 = v[idx5] + v[idx6];
        double sum4 = v[idx7] + v[idx8];
        slow[i] = 8.0/(sum1 + sum2 + sum3 + sum4);
    }
}

__global__ void __launch_bounds__(256)
cell_velocity_to_slowness3d(int nx, int ny, int nz,
                            const double *__restrict__ v, int ordering,
                            double *__restrict__ slow)
{
    const int64_t ncx = nx - 1, ncy = ny - 1, ncz = nz - 1;
    const int64_t nCell = ncx*ncy*ncz;
    for (int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x; i < nCell;
         i += (int64_t) gridDim.x*blockDim.x)
    {
        int64_t src = i;
        if (ordering != 0)
        {
            int64_t iz = i/(ncx*ncy);
            int64_t rem = i - iz*ncx*ncy;
            int64_t iy = rem/ncx;
            int64_t ix = rem - iy*ncx;
            src = (ix*ncy + iy)*ncz + iz;
        }
        slow[i] = 1.0/v[src];
    }
}

template <typename TD, typename TSrc>
__global__ void __launch_bounds__(256)
convert_kernel(int64_t n, const TSrc *__restrict__ in, TD *__restrict__ out)
{
    for (int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x*blockDim.x)
    {
        out[i] = (TD) in[i];
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_for(int64_t n, int block, unsigned cap)
{
    int64_t g = (n + block - 1)/block;
    if (g < 1) { g = 1; }
    return (unsigned) (g < cap ? g : cap);
}

template <typename TS>
hipError_t launch_init3d(const Problem3D *dProblems, int nProblems,
                         int64_t maxNodes, TS *T, hipStream_t stream)
{
    dim3 grid(grid_for(maxNodes, 256, 4096), (unsigned) nProblems);
    hipLaunchKernelGGL(init_travel_times3d<TS>, grid, dim3(256), 0, stream,
                       dProblems, T);
    return hipGetLastError();
}

template <typename TS>
hipError_t launch_hyperplane3d(const Problem3D *dProblems, int problem,
                               int sweep, int initialize, int level, int kLo,
                               int nPairs, TS *T, const TS *S,
                               hipStream_t stream)
{
    if (nPairs <= 0) { return hipSuccess; }
    hipLaunchKernelGGL(sweep_hyperplane3d<TS>,
                       dim3((unsigned) ((nPairs + 255)/256)), dim3(256), 0,
                       stream, dProblems, problem, sweep, initialize, level,
                       kLo, nPairs, T, S);
    return hipGetLastError();
}

template <typename TS>
hipError_t launch_columns3d(const Problem3D *dProblems,
                            const ColumnItem *dItems, int nItems, int sweep,
                            int initialize, TS *T, const TS *S,
                            hipStream_t stream)
{
    if (nItems <= 0) { return hipSuccess; }
    constexpr int W = column_waves<TS>();
    hipLaunchKernelGGL((sweep_columns3d<TS, W, 4>),
                       dim3((unsigned) ((nItems + W - 1)/W)), dim3(64*W), 0,
                       stream, dProblems, dItems, nItems, sweep, initialize,
                       T, S);
    return hipGetLastError();
}

hipError_t launch_nodal_to_cell3d(int nx, int ny, int nz, const double *v,
                                  int ordering, double *slow,
                                  hipStream_t stream)
{
    int64_t nCell = (int64_t) (nx - 1)*(ny - 1)*(nz - 1);
    hipLaunchKernelGGL(nodal_to_cell3d, dim3(grid_for(nCell, 256, 8192)),
                       dim3(256), 0, stream, nx, ny, nz, v, ordering, slow);
    return hipGetLastError();
}

hipError_t launch_cell_velocity_to_slowness3d(int nx, int ny, int nz,
                                              const double *v, int ordering,
                                              double *slow,
                                              hipStream_t stream)
{
    int64_t nCell = (int64_t) (nx - 1)*(ny - 1)*(nz - 1);
    hipLaunchKernelGGL(cell_velocity_to_slowness3d,
                       dim3(grid_for(nCell, 256, 8192)), dim3(256), 0, stream,
                       nx, ny, nz, v, ordering, slow);
    return hipGetLastError();
}

hipError_t launch_convert(int64_t n, const double *in, float *out,
                          hipStream_t stream)
{
    hipLaunchKernelGGL((convert_kernel<float, double>),
                       dim3(grid_for(n, 256, 16384)), dim3(256), 0, stream, n,
                       in, out);
    return hipGetLastError();
}

hipError_t launch_convert(int64_t n, const float *in, double *out,
                          hipStream_t stream)
{
    hipLaunchKernelGGL((convert_kernel<double, float>),
                       dim3(grid_for(n, 256, 16384)), dim3(256), 0, stream, n,
                       in, out);
    return hipGetLastError();
}

template hipError_t launch_init3d<float>(const Problem3D *, int, int64_t,
                                         float *, hipStream_t);
template hipError_t launch_init3d<double>(const Problem3D *, int, int64_t,
                                          double *, hipStream_t);
template hipError_t launch_hyperplane3d<float>(const Problem3D *, int, int,
                                               int, int, int, int, float *,
                                               const float *, hipStream_t);
template hipError_t launch_hyperplane3d<double>(const Problem3D *, int, int,
                                                int, int, int, int, double *,
                                                const double *, hipStream_t);
template hipError_t launch_columns3d<float>(const Problem3D *,
                                            const ColumnItem *, int, int, int,
                                            float *, const float *,
                                            hipStream_t);
template hipError_t launch_columns3d<double>(const Problem3D *,
                                             const ColumnItem *, int, int,
                                             int, double *, const double *,
                                             hipStream_t);

} // namespace ekx
