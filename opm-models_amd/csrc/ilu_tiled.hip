// ilu_tiled.hip -- block ILU(0) factorization and triangular sweeps of the
// 7-point structured-grid pattern as persistent, tile-wavefront kernels, for
// one reservoir or for an ensemble of reservoirs in the same launch.
//
// Semantics are those of the level-scheduled kernels (dune-istl SeqILU<.,.,.,0>,
// natural ordering, istlpreconditionerwrappers.hh:151-190): same rows, same
// per-row operation order.  Only the schedule differs.
//
// Schedule: each grid is cut into BX x BY x BZ tiles, one workgroup of 256
// threads per tile and one thread per cell.  A tile depends on its -x, -y and
// -z neighbour tiles (+x, +y, +z for the backward sweep) of the same
// instance; tiles of all instances are handed out by one atomic ticket in
// order of tile level (TI + TJ + TK), so every dependency of a tile holds a
// smaller ticket.  The kernels are launched cooperatively (all workgroups
// co-resident), hence the tile with the lowest unfinished ticket can always
// proceed and every wait ends; waits are also bounded and report an error
// instead of spinning forever.  Inside a tile the BX+BY+BZ-2 local levels are
// separated by workgroup barriers and exchange values through LDS; values
// across tile faces go through global memory, published with an agent-scope
// release and read once per tile after an agent-scope acquire.
#include <algorithm>
#include <cstring>

#include "ensemble.h"
#include "ilu_tiled.h"

namespace omodels {

namespace {

constexpr int kTileThreads = 256;
constexpr long kSpinLimit = 1L << 22;   // a few seconds of polling

// Polling uses relaxed agent-scope loads (no cache maintenance per poll); the
// single acquire fence after the last dependency is seen invalidates stale
// cache lines once per tile.
__device__ __forceinline__ int loadRelaxed(const int* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void storeRelease(int* p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

struct CellCtx {
    int li, lj, lk;        // local coordinates in the tile
    int i, j, k;           // global coordinates
    long I;
    bool valid;
};

__device__ __forceinline__ void tileOf(const TileGeom& g, int tile, int& TI, int& TJ, int& TK)
{
    TI = tile % g.TX;
    TJ = (tile / g.TX) % g.TY;
    TK = tile / (g.TX * g.TY);
}

__device__ __forceinline__ CellCtx cellOf(const TileGeom& g, int TI, int TJ, int TK)
{
    CellCtx c;
    const int t = threadIdx.x;
    c.li = t % g.BX;
    c.lj = (t / g.BX) % g.BY;
    c.lk = t / (g.BX * g.BY);
    c.i = TI * g.BX + c.li;
    c.j = TJ * g.BY + c.lj;
    c.k = TK * g.BZ + c.lk;
    c.valid = c.i < g.nx && c.j < g.ny && c.k < g.nz && c.lk < g.BZ;
    c.I = (long)c.i + (long)g.nx * ((long)c.j + (long)g.ny * c.k);
    return c;
}

// wait for the dependency tiles of (TI,TJ,TK): dir = -1 forward, +1 backward
__device__ __forceinline__ void waitDeps(const TileGeom& g, int TI, int TJ, int TK, int dir,
                                         const int* done, int epoch, int* err)
{
    if (threadIdx.x == 0) {
        int deps[3];
        int nd = 0;
        const int ti = TI + dir, tj = TJ + dir, tk = TK + dir;
        if (ti >= 0 && ti < g.TX) deps[nd++] = ti + g.TX * (TJ + g.TY * TK);
        if (tj >= 0 && tj < g.TY) deps[nd++] = TI + g.TX * (tj + g.TY * TK);
        if (tk >= 0 && tk < g.TZ) deps[nd++] = TI + g.TX * (TJ + g.TY * tk);
        for (int q = 0; q < nd; ++q) {
            long spins = 0;
            while (loadRelaxed(done + deps[q]) != epoch) {
                if (++spins > kSpinLimit) { atomicOr(err, 1); break; }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

__device__ __forceinline__ int nextTicket(unsigned long long* counter, long long base, int* sTicket)
{
    if (threadIdx.x == 0) *sTicket = (int)((long long)atomicAdd(counter, 1ULL) - base);
    __syncthreads();
    const int t = *sTicket;
    __syncthreads();
    return t;
}

__device__ __forceinline__ void publish(int* done, int tile, int epoch)
{
    __syncthreads();
    if (threadIdx.x == 0) storeRelease(done + tile, epoch);
}

template <class S>
__device__ __forceinline__ void matvecSub(const S* blk, const S* x, S rhs[3])
{
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        S s = blk[e * 3] * x[0];
        s += blk[e * 3 + 1] * x[1];
        s += blk[e * 3 + 2] * x[2];
        rhs[e] += S(-1) * s;
    }
}

__device__ __forceinline__ bool isActive(const IluInst& I) { return I.active == nullptr || *I.active != 0; }

}  // namespace

// ---------------------------------------------------------------------------
// Forward sweep: v = L^-1 d (unit lower), lower blocks in column order -z,-y,-x.
template <class S>
__global__ __launch_bounds__(kTileThreads) void k_ilu_fwd_set(const IluInst* __restrict__ insts,
    const TileTask* __restrict__ order, int T, unsigned long long* counter, long long base, int epoch, int* err)
{
    __shared__ S lv[kTileThreads * 3];
    __shared__ int sTicket;
    for (;;) {
        const int t = nextTicket(counter, base, &sTicket);
        if (t < 0 || t >= T) break;
        const TileTask task = order[t];
        const IluInst I = insts[task.inst];
        if (!isActive(I)) { publish(I.done, task.tile, epoch); continue; }
        const TileGeom& g = I.g;
        const S* __restrict__ lu = (const S*)I.lu;
        const S* __restrict__ d = (const S*)I.d;
        S* __restrict__ v = (S*)I.v;
        const long nxy = (long)g.nx * g.ny;
        int TI, TJ, TK;
        tileOf(g, task.tile, TI, TJ, TK);
        const CellCtx c = cellOf(g, TI, TJ, TK);
        // prefetch the row: rhs and the lower blocks (independent of the sweep)
        S rhs[3] = {S(0), S(0), S(0)};
        S Bz[9], By[9], Bx[9];
        const bool hz = c.k > 0, hy = c.j > 0, hx = c.i > 0;
        if (c.valid) {
#pragma unroll
            for (int e = 0; e < 3; ++e) rhs[e] = d[3 * c.I + e];
            long o = I.rp[c.I];
            if (hz) { for (int q = 0; q < 9; ++q) Bz[q] = lu[9 * o + q]; ++o; }
            if (hy) { for (int q = 0; q < 9; ++q) By[q] = lu[9 * o + q]; ++o; }
            if (hx) { for (int q = 0; q < 9; ++q) Bx[q] = lu[9 * o + q]; }
        }
        waitDeps(g, TI, TJ, TK, -1, I.done, epoch, err);
        // values across the tile faces, all loaded in one round trip
        S xz[3], xy[3], xx[3];
        const bool fz = c.valid && hz && c.lk == 0, fy = c.valid && hy && c.lj == 0, fx = c.valid && hx && c.li == 0;
        if (fz) { for (int e = 0; e < 3; ++e) xz[e] = v[3 * (c.I - nxy) + e]; }
        if (fy) { for (int e = 0; e < 3; ++e) xy[e] = v[3 * (c.I - g.nx) + e]; }
        if (fx) { for (int e = 0; e < 3; ++e) xx[e] = v[3 * (c.I - 1) + e]; }
        const int myLev = c.li + c.lj + c.lk;
        const int nLev = g.BX + g.BY + g.BZ - 2;
        for (int lev = 0; lev <= nLev; ++lev) {
            if (c.valid && myLev == lev) {
                S x[3];
                if (hz) {
                    if (!fz) { for (int e = 0; e < 3; ++e) x[e] = lv[3 * (threadIdx.x - g.BX * g.BY) + e]; }
                    else { for (int e = 0; e < 3; ++e) x[e] = xz[e]; }
                    matvecSub(Bz, x, rhs);
                }
                if (hy) {
                    if (!fy) { for (int e = 0; e < 3; ++e) x[e] = lv[3 * (threadIdx.x - g.BX) + e]; }
                    else { for (int e = 0; e < 3; ++e) x[e] = xy[e]; }
                    matvecSub(By, x, rhs);
                }
                if (hx) {
                    if (!fx) { for (int e = 0; e < 3; ++e) x[e] = lv[3 * (threadIdx.x - 1) + e]; }
                    else { for (int e = 0; e < 3; ++e) x[e] = xx[e]; }
                    matvecSub(Bx, x, rhs);
                }
#pragma unroll
                for (int e = 0; e < 3; ++e) { lv[3 * threadIdx.x + e] = rhs[e]; v[3 * c.I + e] = rhs[e]; }
            }
            __syncthreads();
        }
        publish(I.done, task.tile, epoch);
    }
}

// Backward sweep: v = U^-1 v with the inverted (FP64) pivots; upper blocks +x,+y,+z.
template <class S>
__global__ __launch_bounds__(kTileThreads) void k_ilu_bwd_set(const IluInst* __restrict__ insts,
    const TileTask* __restrict__ order, int T, unsigned long long* counter, long long base, int epoch, int* err)
{
    __shared__ S lv[kTileThreads * 3];
    __shared__ int sTicket;
    for (;;) {
        const int t = nextTicket(counter, base, &sTicket);
        if (t < 0 || t >= T) break;
        const TileTask task = order[T - 1 - t];
        const IluInst I = insts[task.inst];
        if (!isActive(I)) { publish(I.done, task.tile, epoch); continue; }
        const TileGeom& g = I.g;
        const S* __restrict__ lu = (const S*)I.lu;
        const double* __restrict__ dinv = I.dinv;
        S* __restrict__ v = (S*)I.v;
        const long nxy = (long)g.nx * g.ny;
        int TI, TJ, TK;
        tileOf(g, task.tile, TI, TJ, TK);
        const CellCtx c = cellOf(g, TI, TJ, TK);
        S rhs[3] = {S(0), S(0), S(0)};
        S Bx[9], By[9], Bz[9];
        double D[9];
        const bool hx = c.i + 1 < g.nx, hy = c.j + 1 < g.ny, hz = c.k + 1 < g.nz;
        if (c.valid) {
#pragma unroll
            for (int e = 0; e < 3; ++e) rhs[e] = v[3 * c.I + e];
            long o = I.diag[c.I] + 1;
            if (hx) { for (int q = 0; q < 9; ++q) Bx[q] = lu[9 * o + q]; ++o; }
            if (hy) { for (int q = 0; q < 9; ++q) By[q] = lu[9 * o + q]; ++o; }
            if (hz) { for (int q = 0; q < 9; ++q) Bz[q] = lu[9 * o + q]; }
#pragma unroll
            for (int q = 0; q < 9; ++q) D[q] = dinv[9 * c.I + q];
        }
        waitDeps(g, TI, TJ, TK, +1, I.done, epoch, err);
        S xx[3], xy[3], xz[3];
        const bool fx = c.valid && hx && c.li + 1 == g.BX, fy = c.valid && hy && c.lj + 1 == g.BY;
        const bool fz = c.valid && hz && c.lk + 1 == g.BZ;
        if (fx) { for (int e = 0; e < 3; ++e) xx[e] = v[3 * (c.I + 1) + e]; }
        if (fy) { for (int e = 0; e < 3; ++e) xy[e] = v[3 * (c.I + g.nx) + e]; }
        if (fz) { for (int e = 0; e < 3; ++e) xz[e] = v[3 * (c.I + nxy) + e]; }
        const int myLev = (g.BX - 1 - c.li) + (g.BY - 1 - c.lj) + (g.BZ - 1 - c.lk);
        const int nLev = g.BX + g.BY + g.BZ - 2;
        for (int lev = 0; lev <= nLev; ++lev) {
            if (c.valid && myLev == lev) {
                S x[3];
                if (hx) {
                    if (!fx) { for (int e = 0; e < 3; ++e) x[e] = lv[3 * (threadIdx.x + 1) + e]; }
                    else { for (int e = 0; e < 3; ++e) x[e] = xx[e]; }
                    matvecSub(Bx, x, rhs);
                }
                if (hy) {
                    if (!fy) { for (int e = 0; e < 3; ++e) x[e] = lv[3 * (threadIdx.x + g.BX) + e]; }
                    else { for (int e = 0; e < 3; ++e) x[e] = xy[e]; }
                    matvecSub(By, x, rhs);
                }
                if (hz) {
                    if (!fz) { for (int e = 0; e < 3; ++e) x[e] = lv[3 * (threadIdx.x + g.BX * g.BY) + e]; }
                    else { for (int e = 0; e < 3; ++e) x[e] = xz[e]; }
                    matvecSub(Bz, x, rhs);
                }
                S out[3];
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    double s = D[e * 3] * (double)rhs[0];
                    s += D[e * 3 + 1] * (double)rhs[1];
                    s += D[e * 3 + 2] * (double)rhs[2];
                    out[e] = S(0) + S(1) * (S)s;
                }
#pragma unroll
                for (int e = 0; e < 3; ++e) { lv[3 * threadIdx.x + e] = out[e]; v[3 * c.I + e] = out[e]; }
            }
            __syncthreads();
        }
        publish(I.done, task.tile, epoch);
    }
}

// ---------------------------------------------------------------------------
// Factorization: for the 7-point pattern ILU(0) only changes the lower blocks
// (L_IJ = A_IJ A_JJ^-1) and the pivot (A_II -= L_IJ A_JI for J = -z,-y,-x in
// that order), exactly the updates of bilu0_decomposition on this pattern
// (two adjacent cells of a structured grid have no common neighbour, so no
// other entry of the pattern is touched).  Inverted pivots are also written
// row-wise to dinv.
__device__ __forceinline__ void mat3mul(const double* A, const double* B, double* C)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double s = A[i * 3] * B[j];
            s += A[i * 3 + 1] * B[3 + j];
            s += A[i * 3 + 2] * B[6 + j];
            C[i * 3 + j] = s;
        }
}

__device__ __forceinline__ bool inv3(double* a)
{
    double c00 = a[4] * a[8] - a[5] * a[7];
    double c01 = a[5] * a[6] - a[3] * a[8];
    double c02 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (!(det != 0.0) || !isfinite(det)) return false;
    double r = 1.0 / det;
    double c10 = a[2] * a[7] - a[1] * a[8];
    double c11 = a[0] * a[8] - a[2] * a[6];
    double c12 = a[1] * a[6] - a[0] * a[7];
    double c20 = a[1] * a[5] - a[2] * a[4];
    double c21 = a[2] * a[3] - a[0] * a[5];
    double c22 = a[0] * a[4] - a[1] * a[3];
    a[0] = c00 * r; a[1] = c10 * r; a[2] = c20 * r;
    a[3] = c01 * r; a[4] = c11 * r; a[5] = c21 * r;
    a[6] = c02 * r; a[7] = c12 * r; a[8] = c22 * r;
    return true;
}

__global__ __launch_bounds__(kTileThreads) void k_ilu_factor_set(const IluInst* __restrict__ insts,
    const TileTask* __restrict__ order, int T, unsigned long long* counter, long long base, int epoch, int* err,
    int* fail)
{
    __shared__ double ld[kTileThreads * 9];
    __shared__ int sTicket;
    for (;;) {
        const int t = nextTicket(counter, base, &sTicket);
        if (t < 0 || t >= T) break;
        const TileTask task = order[t];
        const IluInst I = insts[task.inst];
        if (!isActive(I)) { publish(I.done, task.tile, epoch); continue; }
        const TileGeom& g = I.g;
        double* __restrict__ lu = (double*)I.luw;
        double* __restrict__ dinv = I.dinvw;
        const long nxy = (long)g.nx * g.ny;
        int TI, TJ, TK;
        tileOf(g, task.tile, TI, TJ, TK);
        const CellCtx c = cellOf(g, TI, TJ, TK);
        const bool hz = c.k > 0, hy = c.j > 0, hx = c.i > 0;
        double D[9], Az[9], Ay[9], Ax[9], Uz[9], Uy[9], Ux[9];
        long oz = 0, oy = 0, ox = 0, od = 0;
        if (c.valid) {
            long o = I.rp[c.I];
            if (hz) { oz = o++; }
            if (hy) { oy = o++; }
            if (hx) { ox = o++; }
            od = I.diag[c.I];
            for (int q = 0; q < 9; ++q) D[q] = lu[9 * od + q];
            // A_JI: the block of neighbour row J that points back to I
            if (hz) {
                const long J = c.I - nxy;
                const long oJ = I.diag[J] + 1 + (c.i + 1 < g.nx) + (c.j + 1 < g.ny);
                for (int q = 0; q < 9; ++q) { Az[q] = lu[9 * oz + q]; Uz[q] = lu[9 * oJ + q]; }
            }
            if (hy) {
                const long J = c.I - g.nx;
                const long oJ = I.diag[J] + 1 + (c.i + 1 < g.nx);
                for (int q = 0; q < 9; ++q) { Ay[q] = lu[9 * oy + q]; Uy[q] = lu[9 * oJ + q]; }
            }
            if (hx) {
                const long J = c.I - 1;
                const long oJ = I.diag[J] + 1;
                for (int q = 0; q < 9; ++q) { Ax[q] = lu[9 * ox + q]; Ux[q] = lu[9 * oJ + q]; }
            }
        }
        waitDeps(g, TI, TJ, TK, -1, I.done, epoch, err);
        // pivots of the neighbours across the tile faces, one round trip
        const bool fz = c.valid && hz && c.lk == 0, fy = c.valid && hy && c.lj == 0, fx = c.valid && hx && c.li == 0;
        double Pz[9], Py[9], Px[9];
        if (fz) { for (int q = 0; q < 9; ++q) Pz[q] = dinv[9 * (c.I - nxy) + q]; }
        if (fy) { for (int q = 0; q < 9; ++q) Py[q] = dinv[9 * (c.I - g.nx) + q]; }
        if (fx) { for (int q = 0; q < 9; ++q) Px[q] = dinv[9 * (c.I - 1) + q]; }
        const int myLev = c.li + c.lj + c.lk;
        const int nLev = g.BX + g.BY + g.BZ - 2;
        for (int lev = 0; lev <= nLev; ++lev) {
            if (c.valid && myLev == lev) {
                double Dj[9], L[9], Tm[9];
                if (hz) {
                    if (!fz) { for (int q = 0; q < 9; ++q) Dj[q] = ld[9 * (threadIdx.x - g.BX * g.BY) + q]; }
                    else { for (int q = 0; q < 9; ++q) Dj[q] = Pz[q]; }
                    mat3mul(Az, Dj, L);
                    for (int q = 0; q < 9; ++q) lu[9 * oz + q] = L[q];
                    mat3mul(L, Uz, Tm);
                    for (int q = 0; q < 9; ++q) D[q] -= Tm[q];
                }
                if (hy) {
                    if (!fy) { for (int q = 0; q < 9; ++q) Dj[q] = ld[9 * (threadIdx.x - g.BX) + q]; }
                    else { for (int q = 0; q < 9; ++q) Dj[q] = Py[q]; }
                    mat3mul(Ay, Dj, L);
                    for (int q = 0; q < 9; ++q) lu[9 * oy + q] = L[q];
                    mat3mul(L, Uy, Tm);
                    for (int q = 0; q < 9; ++q) D[q] -= Tm[q];
                }
                if (hx) {
                    if (!fx) { for (int q = 0; q < 9; ++q) Dj[q] = ld[9 * (threadIdx.x - 1) + q]; }
                    else { for (int q = 0; q < 9; ++q) Dj[q] = Px[q]; }
                    mat3mul(Ax, Dj, L);
                    for (int q = 0; q < 9; ++q) lu[9 * ox + q] = L[q];
                    mat3mul(L, Ux, Tm);
                    for (int q = 0; q < 9; ++q) D[q] -= Tm[q];
                }
                if (!inv3(D)) atomicOr(fail, 2);
                for (int q = 0; q < 9; ++q) {
                    ld[9 * threadIdx.x + q] = D[q];
                    lu[9 * od + q] = D[q];
                    dinv[9 * c.I + q] = D[q];
                }
            }
            __syncthreads();
        }
        publish(I.done, task.tile, epoch);
    }
}

// ---------------------------------------------------------------------------
TileGeom makeTileGeom(int nx, int ny, int nz)
{
    // tile shape: 256 cells, as cubic as the grid allows
    int bz = nz >= 4 ? 4 : (nz >= 2 ? 2 : 1);
    int rest = kTileThreads / bz;
    int bx = 1;
    while (bx * bx < rest) bx *= 2;
    int by = rest / bx;
    if (nx < bx) { bx = 1; while (bx * 2 <= nx) bx *= 2; by = rest / bx; }
    if (ny < by) { by = 1; while (by * 2 <= ny) by *= 2; }
    TileGeom g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.BX = bx; g.BY = by; g.BZ = bz;
    g.TX = (nx + bx - 1) / bx; g.TY = (ny + by - 1) / by; g.TZ = (nz + bz - 1) / bz;
    g.T = g.TX * g.TY * g.TZ;
    return g;
}

TiledIluSet::TiledIluSet(hipStream_t s)
{
    counter_.alloc(1);
    OM_HIP(hipMemsetAsync(counter_.p, 0, sizeof(unsigned long long), s));
    err_.alloc(1);
    OM_HIP(hipMemsetAsync(err_.p, 0, sizeof(int), s));
    int dev = 0;
    OM_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    OM_HIP(hipGetDeviceProperties(&prop, dev));
    const void* ks[] = {(const void*)k_ilu_factor_set, (const void*)k_ilu_fwd_set<double>,
                        (const void*)k_ilu_bwd_set<double>, (const void*)k_ilu_fwd_set<float>,
                        (const void*)k_ilu_bwd_set<float>};
    int perCU = 1 << 20;
    for (const void* k : ks) {
        int n = 0;
        OM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, kTileThreads, 0));
        perCU = std::min(perCU, n);
    }
    if (perCU < 1) throw std::runtime_error("tiled ILU kernels cannot be resident");
    maxResident_ = perCU * prop.multiProcessorCount;
}

void TiledIluSet::setInstances(const std::vector<TileGeom>& geoms, hipStream_t s)
{
    // tickets: all tiles of all instances ordered by tile level, then instance
    std::vector<std::pair<int, TileTask>> tasks;
    tileOff_.assign(geoms.size() + 1, 0);
    for (size_t b = 0; b < geoms.size(); ++b) {
        const TileGeom& g = geoms[b];
        tileOff_[b + 1] = tileOff_[b] + g.T;
        for (int t = 0; t < g.T; ++t) {
            const int TI = t % g.TX, TJ = (t / g.TX) % g.TY, TK = t / (g.TX * g.TY);
            tasks.push_back({TI + TJ + TK, TileTask{(int)b, t}});
        }
    }
    std::stable_sort(tasks.begin(), tasks.end(),
                     [](const std::pair<int, TileTask>& a, const std::pair<int, TileTask>& b) { return a.first < b.first; });
    std::vector<TileTask> ord(tasks.size());
    for (size_t k = 0; k < tasks.size(); ++k) ord[k] = tasks[k].second;
    OM_HIP(hipStreamSynchronize(s));
    T_ = (int)ord.size();
    order_.upload(ord.data(), ord.size(), s);
    done_.alloc(std::max(T_, 1));
    OM_HIP(hipMemsetAsync(done_.p, 0, std::max(T_, 1) * sizeof(int), s));
    // flags restart from zero: restart the epoch so that no stale value matches
    epoch_ = 0;
    grid_ = std::min(T_, maxResident_);
    dcache_.clear();
}

const IluInst* TiledIluSet::upload(const std::vector<IluInst>& insts, hipStream_t s)
{
    std::vector<IluInst> v(insts);
    for (size_t b = 0; b < v.size(); ++b) v[b].done = done_.p + tileOff_[b];
    // descriptor arrays stay resident: a solver alternates between a few
    // (input, output) vector pairs, each gets its own device copy once
    for (auto& e : dcache_)
        if (e.first.size() == v.size() && std::memcmp(e.first.data(), v.data(), v.size() * sizeof(IluInst)) == 0)
            return e.second->p;
    if (dcache_.size() >= 16) {
        OM_HIP(hipStreamSynchronize(s));
        dcache_.clear();
    }
    auto buf = std::make_unique<DevBuf<IluInst>>();
    buf->upload(v.data(), v.size(), s);
    dcache_.emplace_back(v, std::move(buf));
    return dcache_.back().second->p;
}

void TiledIluSet::checkError(hipStream_t s)
{
    int e = 0;
    OM_HIP(hipMemcpyAsync(&e, err_.p, sizeof(int), hipMemcpyDeviceToHost, s));
    OM_HIP(hipStreamSynchronize(s));
    if (e) throw std::runtime_error("tiled ILU sweep: dependency wait exceeded its bound");
}

void TiledIluSet::factor(const std::vector<IluInst>& insts, int* fail, hipStream_t s)
{
    if (T_ == 0) return;
    const IluInst* di = upload(insts, s);
    long long b = ticketBase_;
    ticketBase_ += T_ + grid_;   // every launch draws exactly T + grid tickets
    int ep = ++epoch_;
    const TileTask* od = order_.p;
    int T = T_;
    unsigned long long* ct = counter_.p;
    int* er = err_.p;
    void* args[] = {(void*)&di, (void*)&od, &T, &ct, &b, &ep, &er, (void*)&fail};
    OM_HIP(hipLaunchCooperativeKernel((const void*)k_ilu_factor_set, dim3(grid_), dim3(kTileThreads), args, 0, s));
}

template <class S>
void TiledIluSet::apply(const std::vector<IluInst>& insts, hipStream_t s)
{
    if (T_ == 0) return;
    const IluInst* di = upload(insts, s);
    const TileTask* od = order_.p;
    int T = T_;
    unsigned long long* ct = counter_.p;
    int* er = err_.p;
    {
        long long b = ticketBase_;
        ticketBase_ += T_ + grid_;   // every launch draws exactly T + grid tickets
        int ep = ++epoch_;
        void* args[] = {(void*)&di, (void*)&od, &T, &ct, &b, &ep, &er};
        OM_HIP(hipLaunchCooperativeKernel((const void*)k_ilu_fwd_set<S>, dim3(grid_), dim3(kTileThreads), args, 0, s));
    }
    {
        long long b = ticketBase_;
        ticketBase_ += T_ + grid_;   // every launch draws exactly T + grid tickets
        int ep = ++epoch_;
        void* args[] = {(void*)&di, (void*)&od, &T, &ct, &b, &ep, &er};
        OM_HIP(hipLaunchCooperativeKernel((const void*)k_ilu_bwd_set<S>, dim3(grid_), dim3(kTileThreads), args, 0, s));
    }
}
template void TiledIluSet::apply<double>(const std::vector<IluInst>&, hipStream_t);
template void TiledIluSet::apply<float>(const std::vector<IluInst>&, hipStream_t);

// ---------------------------------------------------------------------------
// single reservoir: an ensemble of one
TiledIlu::TiledIlu(int nx, int ny, int nz, hipStream_t s) : set_(new TiledIluSet(s))
{
    g = makeTileGeom(nx, ny, nz);
    set_->setInstances({g}, s);
}

TiledIlu::~TiledIlu() { delete set_; }

void TiledIlu::checkError(hipStream_t s) { set_->checkError(s); }

void TiledIlu::factor(const int* rp, const int* diag, double* lu, double* dinvOut, int* fail, hipStream_t s)
{
    IluInst I{};
    I.g = g; I.rp = rp; I.diag = diag; I.lu = lu; I.luw = lu; I.dinv = dinvOut; I.dinvw = dinvOut;
    I.active = nullptr;
    set_->factor({I}, fail, s);
}

template <class S, class P>
void TiledIlu::apply(const int* rp, const int* diag, const S* lu, const P* dinv, const S* d, S* v, hipStream_t s)
{
    static_assert(sizeof(P) == 8, "pivots are kept in FP64");
    IluInst I{};
    I.g = g; I.rp = rp; I.diag = diag; I.lu = lu; I.dinv = (const double*)dinv; I.d = d; I.v = v;
    I.active = nullptr;
    set_->apply<S>({I}, s);
}
template void TiledIlu::apply<double, double>(const int*, const int*, const double*, const double*, const double*,
                                              double*, hipStream_t);
template void TiledIlu::apply<float, double>(const int*, const int*, const float*, const double*, const float*,
                                             float*, hipStream_t);

}  // namespace omodels
