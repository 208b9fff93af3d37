// linalg_kernels.hip -- block-CSR linear algebra for the BiCGStab backend.
//
//   SpMV                     Dune BCRSMatrix::mv as called by OverlappingOperator::apply
//                            (overlappingoperator.hh:58-62)
//   block ILU(0) factor/apply dune-istl SeqILU<.,.,.,0> via PreconditionerWrapperILU
//                            (istlpreconditionerwrappers.hh:151-190), level-scheduled
//   block Jacobi             PreconditionerWrapperJacobi (istlpreconditionerwrappers.hh:143)
//   BiCGStab vector steps    BiCGStabSolver::apply (bicgstabsolver.hh:120-330)
//   CombinedCriterion norms  combinedcriterion.hh:54-236
//
// Reductions use a fixed number of partial sums and a fixed tree, so results
// are reproducible run to run; BiCGStab scalars stay in device memory.
#include <hipcub/hipcub.hpp>

#include "exact_sum.h"
#include "omodels_internal.h"

namespace omodels {

// ---------------------------------------------------------------------------
// small dense block helpers (row-major, BS = 2 or 3)
template <int BS, class S>
__device__ __forceinline__ void bMatmul(const S* A, const S* B, S* C)
{
#pragma unroll
    for (int i = 0; i < BS; ++i)
#pragma unroll
        for (int j = 0; j < BS; ++j) {
            S s = A[i * BS] * B[j];
#pragma unroll
            for (int k = 1; k < BS; ++k) s += A[i * BS + k] * B[k * BS + j];
            C[i * BS + j] = s;
        }
}

template <int BS, class S>
__device__ __forceinline__ bool bInvert(S* a)
{
    if (BS == 2) {
        S det = a[0] * a[3] - a[1] * a[2];
        if (!(det != S(0)) || !isfinite((double)det)) return false;
        S r = S(1) / det;
        S i00 = a[3] * r, i01 = -a[1] * r, i10 = -a[2] * r, i11 = a[0] * r;
        a[0] = i00; a[1] = i01; a[2] = i10; a[3] = i11;
        return true;
    } else {
        S c00 = a[4] * a[8] - a[5] * a[7];
        S c01 = a[5] * a[6] - a[3] * a[8];
        S c02 = a[3] * a[7] - a[4] * a[6];
        S det = a[0] * c00 + a[1] * c01 + a[2] * c02;
        if (!(det != S(0)) || !isfinite((double)det)) return false;
        S r = S(1) / det;
        S c10 = a[2] * a[7] - a[1] * a[8];
        S c11 = a[0] * a[8] - a[2] * a[6];
        S c12 = a[1] * a[6] - a[0] * a[7];
        S c20 = a[1] * a[5] - a[2] * a[4];
        S c21 = a[2] * a[3] - a[0] * a[5];
        S c22 = a[0] * a[4] - a[1] * a[3];
        a[0] = c00 * r; a[1] = c10 * r; a[2] = c20 * r;
        a[3] = c01 * r; a[4] = c11 * r; a[5] = c21 * r;
        a[6] = c02 * r; a[7] = c12 * r; a[8] = c22 * r;
        return true;
    }
}

// ---------------------------------------------------------------------------
// SpMV: one thread per block row.
template <int BS, class S>
__global__ __launch_bounds__(256) void k_spmv(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                              const S* __restrict__ a, const S* __restrict__ x,
                                              S* __restrict__ y, const int* __restrict__ run)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (run && !*run)) return;
    S acc[BS];
#pragma unroll
    for (int e = 0; e < BS; ++e) acc[e] = S(0);
    const int k1 = rp[i + 1];
    for (int k = rp[i]; k < k1; ++k) {
        const int c = ci[k];
        const S* blk = a + (long)BS * BS * k;
        S xv[BS];
#pragma unroll
        for (int q = 0; q < BS; ++q) xv[q] = x[(long)BS * c + q];
#pragma unroll
        for (int e = 0; e < BS; ++e) {
            S s = blk[e * BS] * xv[0];
#pragma unroll
            for (int q = 1; q < BS; ++q) s += blk[e * BS + q] * xv[q];
            acc[e] += s;
        }
    }
#pragma unroll
    for (int e = 0; e < BS; ++e) y[(long)BS * i + e] = acc[e];
}

// Workgroup w runs on XCD w % 8 (round-robin dispatch); renumber the
// workgroups so that each XCD walks one contiguous range of rows and the
// x values of neighbouring rows (+-1, +-nx, +-nx*ny) are reused in its own L2.
__device__ __forceinline__ int xcdContiguousBlock(int w, int G)
{
    const int per = G / 8, rem = G % 8, xcd = w % 8, k = w / 8;
    return xcd * per + (xcd < rem ? xcd : rem) + k;
}

// SpMV for 3x3 blocks: three lanes per block row, lane e forms row e of each
// block product, so a wave reads each row's blocks as contiguous 24-byte
// pieces; per row the operation order is that of k_spmv.
template <class S>
__global__ __launch_bounds__(256) void k_spmv3(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                               const S* __restrict__ a, const S* __restrict__ x,
                                               S* __restrict__ y, const int* __restrict__ run)
{
    if (run && !*run) return;
    const long g = (long)xcdContiguousBlock(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int i = (int)(g / 3), e = (int)(g - 3L * (g / 3));
    if (i >= n) return;
    // a 7-point row has at most 7 blocks: the fixed bound lets every load of
    // the row be issued before the dependent sums
    const int k0 = rp[i], nb = rp[i + 1] - k0;
    S v[7][3], xv[7][3];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        if (j < nb) {
            const long c = ci[k0 + j];
            const S* row = a + 9L * (k0 + j) + 3 * e;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                v[j][q] = row[q];
                xv[j][q] = x[3 * c + q];
            }
        }
    }
    S acc = S(0);
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        if (j < nb) {
            S s = v[j][0] * xv[j][0];
            s += v[j][1] * xv[j][1];
            s += v[j][2] * xv[j][2];
            acc += s;
        }
    }
    for (int k = k0 + 7; k < k0 + nb; ++k) {   // longer rows (other patterns), same order
        const long c = ci[k];
        const S* row = a + 9L * k + 3 * e;
        S s = row[0] * x[3 * c];
        s += row[1] * x[3 * c + 1];
        s += row[2] * x[3 * c + 2];
        acc += s;
    }
    y[3L * i + e] = acc;
}

template <class S>
void launchSpmv(int n, int bs, const int* rp, const int* ci, const S* a, const S* x, S* y, hipStream_t st,
                const int* run)
{
    if (bs == 3) {
        const long lanes = 3L * n;
        hipLaunchKernelGGL((k_spmv3<S>), dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, st, n, rp, ci, a, x, y,
                           run);
    } else {
        dim3 g((n + 255) / 256), b(256);
        hipLaunchKernelGGL((k_spmv<2, S>), g, b, 0, st, n, rp, ci, a, x, y, run);
    }
}
// ---------------------------------------------------------------------------
// The structured layout of a 7-point 3x3 BCRS matrix on an nx x ny x nz grid:
// D[slot][I][9] with slots -z, -y, -x, I, +x, +y, +z (the sorted row order),
// zero blocks where a neighbour is absent.  The blocks of one slot of
// consecutive cells are consecutive, so the sweeps and the operator read
// whole cache lines.
__device__ __forceinline__ bool diaPresent(int slot, int i, int j, int k, int nx, int ny, int nz)
{
    switch (slot) {
    case 0: return k > 0;
    case 1: return j > 0;
    case 2: return i > 0;
    case 3: return true;
    case 4: return i + 1 < nx;
    case 5: return j + 1 < ny;
    default: return k + 1 < nz;
    }
}
__device__ __forceinline__ long diaNeighbour(int slot, long I, int nx, long nxy)
{
    switch (slot) {
    case 0: return I - nxy;
    case 1: return I - nx;
    case 2: return I - 1;
    case 3: return I;
    case 4: return I + 1;
    case 5: return I + nx;
    default: return I + nxy;
    }
}

template <class S>
__global__ __launch_bounds__(256) void k_to_dia(int nx, int ny, int nz, const int* __restrict__ rp,
                                                const double* __restrict__ A, S* __restrict__ D)
{
    const long N = (long)nx * ny * nz;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 7 * N) return;
    const int slot = (int)(t / N);
    const long I = t - slot * N;
    const int i = (int)(I % nx), j = (int)((I / nx) % ny), k = (int)(I / ((long)nx * ny));
    S* out = D + 9 * t;
    if (!diaPresent(slot, i, j, k, nx, ny, nz)) {
#pragma unroll
        for (int c = 0; c < 9; ++c) out[c] = S(0);
        return;
    }
    long pos = rp[I];
    for (int q = 0; q < slot; ++q) pos += diaPresent(q, i, j, k, nx, ny, nz) ? 1 : 0;
    const double* in = A + 9 * pos;
#pragma unroll
    for (int c = 0; c < 9; ++c) out[c] = (S)in[c];
}

// y = A x in the structured layout: the arithmetic of k_spmv3 (row sums of
// the present blocks in column order), so the same bits
template <class S>
__global__ __launch_bounds__(256) void k_spmv_dia(int nx, int ny, int nz, const S* __restrict__ D,
                                                  const S* __restrict__ x, S* __restrict__ y,
                                                  const int* __restrict__ run)
{
    if (run && !*run) return;
    const long N = (long)nx * ny * nz;
    const long nxy = (long)nx * ny;
    const long g = (long)xcdContiguousBlock(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const long I = g / 3;
    const int e = (int)(g - 3 * I);
    if (I >= N) return;
    const int i = (int)(I % nx), j = (int)((I / nx) % ny), k = (int)(I / nxy);
    S v[7][3], xv[7][3];
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        if (diaPresent(q, i, j, k, nx, ny, nz)) {
            const long c = diaNeighbour(q, I, nx, nxy);
            const S* row = D + 9 * (q * N + I) + 3 * e;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                v[q][r] = row[r];
                xv[q][r] = x[3 * c + r];
            }
        }
    }
    S acc = S(0);
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        if (diaPresent(q, i, j, k, nx, ny, nz)) {
            S s = v[q][0] * xv[q][0];
            s += v[q][1] * xv[q][1];
            s += v[q][2] * xv[q][2];
            acc += s;
        }
    }
    y[3 * I + e] = acc;
}

template <class S>
void launchToDia(int nx, int ny, int nz, const int* rp, const double* A, S* D, hipStream_t st)
{
    const long n = 7L * nx * ny * nz;
    if (n <= 0) return;
    hipLaunchKernelGGL((k_to_dia<S>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nx, ny, nz, rp, A, D);
}
template void launchToDia<double>(int, int, int, const int*, const double*, double*, hipStream_t);
template void launchToDia<float>(int, int, int, const int*, const double*, float*, hipStream_t);

template <class S>
void launchSpmvDia(int nx, int ny, int nz, const S* D, const S* x, S* y, hipStream_t st, const int* run)
{
    const long lanes = 3L * nx * ny * nz;
    if (lanes <= 0) return;
    hipLaunchKernelGGL((k_spmv_dia<S>), dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, st, nx, ny, nz, D, x, y,
                       run);
}
template void launchSpmvDia<double>(int, int, int, const double*, const double*, double*, hipStream_t, const int*);
template void launchSpmvDia<float>(int, int, int, const float*, const float*, float*, hipStream_t, const int*);

template void launchSpmv<double>(int, int, const int*, const int*, const double*, const double*, double*, hipStream_t,
                                 const int*);
template void launchSpmv<float>(int, int, const int*, const int*, const float*, const float*, float*, hipStream_t,
                                const int*);

// ---------------------------------------------------------------------------
// conversions
template <class S>
__global__ void k_convert(long m, const double* __restrict__ x, S* __restrict__ y)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) y[i] = (S)x[i];
}
template <class S>
void launchCopyConvertD(long m, const double* x, S* y, hipStream_t st)
{
    if (m <= 0) return;
    hipLaunchKernelGGL((k_convert<S>), dim3((m + 255) / 256), dim3(256), 0, st, m, x, y);
}
template void launchCopyConvertD<double>(long, const double*, double*, hipStream_t);
template void launchCopyConvertD<float>(long, const double*, float*, hipStream_t);

// ---------------------------------------------------------------------------
// ILU(0), one launch per level; rows of a level are independent.
template <int BS, class S>
__global__ __launch_bounds__(256) void k_ilu_factor_level(const int* __restrict__ rows, int cnt,
    const int* __restrict__ rp, const int* __restrict__ ci, const int* __restrict__ diag,
    S* __restrict__ lu, int* __restrict__ fail)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int i = rows[t];
    constexpr int BB = BS * BS;
    const int e1 = rp[i + 1];
    int ij = rp[i];
    for (; ij < e1 && ci[ij] < i; ++ij) {
        const int j = ci[ij];
        const int jj = diag[j];
        S Lij[BB], tmp[BB];
        {
            S Aij[BB], Djj[BB];
#pragma unroll
            for (int q = 0; q < BB; ++q) { Aij[q] = lu[(long)BB * ij + q]; Djj[q] = lu[(long)BB * jj + q]; }
            bMatmul<BS, S>(Aij, Djj, Lij);
        }
#pragma unroll
        for (int q = 0; q < BB; ++q) lu[(long)BB * ij + q] = Lij[q];
        int ik = ij + 1, jk = jj + 1;
        const int je = rp[j + 1];
        while (ik < e1 && jk < je) {
            const int cik = ci[ik], cjk = ci[jk];
            if (cik == cjk) {
                S Ajk[BB];
#pragma unroll
                for (int q = 0; q < BB; ++q) Ajk[q] = lu[(long)BB * jk + q];
                bMatmul<BS, S>(Lij, Ajk, tmp);
#pragma unroll
                for (int q = 0; q < BB; ++q) lu[(long)BB * ik + q] -= tmp[q];
                ++ik; ++jk;
            } else if (cik < cjk) ++ik;
            else ++jk;
        }
    }
    if (ij >= e1 || ci[ij] != i) { atomicOr(fail, 1); return; }
    S D[BB];
#pragma unroll
    for (int q = 0; q < BB; ++q) D[q] = lu[(long)BB * ij + q];
    if (!bInvert<BS, S>(D)) atomicOr(fail, 2);
#pragma unroll
    for (int q = 0; q < BB; ++q) lu[(long)BB * ij + q] = D[q];
}

template <class S>
int launchIlu0Factor(int n, int bs, const int* rp, const int* ci, const int* diag, S* lu,
                     const LevelSchedule& L, int* dFail, hipStream_t st)
{
    (void)n;
    for (int l = 0; l < L.nLevels; ++l) {
        const int b = L.ptrHost[l], cnt = L.ptrHost[l + 1] - b;
        dim3 g((cnt + 255) / 256), blk(256);
        if (bs == 3) hipLaunchKernelGGL((k_ilu_factor_level<3, S>), g, blk, 0, st, L.rows.p + b, cnt, rp, ci, diag, lu, dFail);
        else hipLaunchKernelGGL((k_ilu_factor_level<2, S>), g, blk, 0, st, L.rows.p + b, cnt, rp, ci, diag, lu, dFail);
    }
    return 0;
}
template int launchIlu0Factor<double>(int, int, const int*, const int*, const int*, double*, const LevelSchedule&, int*, hipStream_t);
template int launchIlu0Factor<float>(int, int, const int*, const int*, const int*, float*, const LevelSchedule&, int*, hipStream_t);

template <int BS, class S>
__global__ __launch_bounds__(256) void k_ilu_fwd_level(const int* __restrict__ rows, int cnt,
    const int* __restrict__ rp, const int* __restrict__ ci, const int* __restrict__ diag,
    const S* __restrict__ lu, const S* __restrict__ d, S* __restrict__ v)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int i = rows[t];
    constexpr int BB = BS * BS;
    S rhs[BS];
#pragma unroll
    for (int e = 0; e < BS; ++e) rhs[e] = d[(long)BS * i + e];
    const int kd = diag[i];
    for (int k = rp[i]; k < kd; ++k) {
        const int c = ci[k];
        const S* blk = lu + (long)BB * k;
        S xv[BS];
#pragma unroll
        for (int q = 0; q < BS; ++q) xv[q] = v[(long)BS * c + q];
#pragma unroll
        for (int e = 0; e < BS; ++e) {
            S s = blk[e * BS] * xv[0];
#pragma unroll
            for (int q = 1; q < BS; ++q) s += blk[e * BS + q] * xv[q];
            rhs[e] += S(-1) * s;
        }
    }
#pragma unroll
    for (int e = 0; e < BS; ++e) v[(long)BS * i + e] = rhs[e];
}

template <int BS, class S, class P>
__global__ __launch_bounds__(256) void k_ilu_bwd_level(const int* __restrict__ rows, int cnt,
    const int* __restrict__ rp, const int* __restrict__ ci, const int* __restrict__ diag,
    const S* __restrict__ lu, const P* __restrict__ dinv, S* __restrict__ v, S relax)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int i = rows[t];
    constexpr int BB = BS * BS;
    S rhs[BS];
#pragma unroll
    for (int e = 0; e < BS; ++e) rhs[e] = v[(long)BS * i + e];
    const int kd = diag[i], k1 = rp[i + 1];
    for (int k = kd + 1; k < k1; ++k) {
        const int c = ci[k];
        const S* blk = lu + (long)BB * k;
        S xv[BS];
#pragma unroll
        for (int q = 0; q < BS; ++q) xv[q] = v[(long)BS * c + q];
#pragma unroll
        for (int e = 0; e < BS; ++e) {
            S s = blk[e * BS] * xv[0];
#pragma unroll
            for (int q = 1; q < BS; ++q) s += blk[e * BS + q] * xv[q];
            rhs[e] += S(-1) * s;
        }
    }
    const P* D = dinv + (long)BB * i;
#pragma unroll
    for (int e = 0; e < BS; ++e) {
        P s = D[e * BS] * (P)rhs[0];
#pragma unroll
        for (int q = 1; q < BS; ++q) s += D[e * BS + q] * (P)rhs[q];
        S o = S(0) + S(1) * (S)s;
        if (relax != S(1)) o *= relax;
        v[(long)BS * i + e] = o;
    }
}

template <class S, class P>
void launchIlu0Apply(int n, int bs, const int* rp, const int* ci, const int* diag, const S* lu,
                     const P* dinv, const S* d, S* v, const LevelSchedule& L, hipStream_t st)
{
    (void)n;
    const S relax = S(1);
    for (int l = 0; l < L.nLevels; ++l) {
        const int b = L.ptrHost[l], cnt = L.ptrHost[l + 1] - b;
        dim3 g((cnt + 255) / 256), blk(256);
        if (bs == 3) hipLaunchKernelGGL((k_ilu_fwd_level<3, S>), g, blk, 0, st, L.rows.p + b, cnt, rp, ci, diag, lu, d, v);
        else hipLaunchKernelGGL((k_ilu_fwd_level<2, S>), g, blk, 0, st, L.rows.p + b, cnt, rp, ci, diag, lu, d, v);
    }
    for (int l = L.nLevels - 1; l >= 0; --l) {
        const int b = L.ptrHost[l], cnt = L.ptrHost[l + 1] - b;
        dim3 g((cnt + 255) / 256), blk(256);
        if (bs == 3) hipLaunchKernelGGL((k_ilu_bwd_level<3, S, P>), g, blk, 0, st, L.rows.p + b, cnt, rp, ci, diag, lu, dinv, v, relax);
        else hipLaunchKernelGGL((k_ilu_bwd_level<2, S, P>), g, blk, 0, st, L.rows.p + b, cnt, rp, ci, diag, lu, dinv, v, relax);
    }
}
template void launchIlu0Apply<double, double>(int, int, const int*, const int*, const int*, const double*, const double*, const double*, double*, const LevelSchedule&, hipStream_t);
template void launchIlu0Apply<float, float>(int, int, const int*, const int*, const int*, const float*, const float*, const float*, float*, const LevelSchedule&, hipStream_t);
template void launchIlu0Apply<float, double>(int, int, const int*, const int*, const int*, const float*, const double*, const float*, float*, const LevelSchedule&, hipStream_t);

// inverted pivots gathered per row (optionally in a wider type)
template <int BS, class S>
__global__ void k_extract_dinv(int n, const int* __restrict__ diag, const S* __restrict__ lu,
                               double* __restrict__ dinv)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    constexpr int BB = BS * BS;
#pragma unroll
    for (int q = 0; q < BB; ++q) dinv[(long)BB * i + q] = (double)lu[(long)BB * diag[i] + q];
}
template <class S>
void launchExtractInvDiag(int n, int bs, const int* diag, const S* lu, double* dinv, hipStream_t st)
{
    dim3 g((n + 255) / 256), b(256);
    if (bs == 3) hipLaunchKernelGGL((k_extract_dinv<3, S>), g, b, 0, st, n, diag, lu, dinv);
    else hipLaunchKernelGGL((k_extract_dinv<2, S>), g, b, 0, st, n, diag, lu, dinv);
}
template void launchExtractInvDiag<double>(int, int, const int*, const double*, double*, hipStream_t);
template void launchExtractInvDiag<float>(int, int, const int*, const float*, double*, hipStream_t);

// ---------------------------------------------------------------------------
// block Jacobi: dinv = inverse of the diagonal blocks; v = w * dinv * d
template <int BS, class S>
__global__ void k_jacobi_invert(int n, const int* __restrict__ diag, const S* __restrict__ a,
                                S* __restrict__ dinv, int* __restrict__ fail)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    constexpr int BB = BS * BS;
    S D[BB];
#pragma unroll
    for (int q = 0; q < BB; ++q) D[q] = a[(long)BB * diag[i] + q];
    if (!bInvert<BS, S>(D)) atomicOr(fail, 2);
#pragma unroll
    for (int q = 0; q < BB; ++q) dinv[(long)BB * i + q] = D[q];
}
template <class S>
void launchJacobiInvert(int n, int bs, const int* diag, const S* a, S* dinv, int* dFail, hipStream_t st)
{
    dim3 g((n + 255) / 256), b(256);
    if (bs == 3) hipLaunchKernelGGL((k_jacobi_invert<3, S>), g, b, 0, st, n, diag, a, dinv, dFail);
    else hipLaunchKernelGGL((k_jacobi_invert<2, S>), g, b, 0, st, n, diag, a, dinv, dFail);
}
template void launchJacobiInvert<double>(int, int, const int*, const double*, double*, int*, hipStream_t);
template void launchJacobiInvert<float>(int, int, const int*, const float*, float*, int*, hipStream_t);

// One block-Jacobi sweep, Dune::SeqJac's step: xout = (1 - w) xin
// + w D^{-1} (d - (L + U) xin), the products in the operation order of the
// other kernels (row sums, then the subtraction); D^{-1} in FP64.
template <class S>
__global__ __launch_bounds__(256) void k_jacobi_sweep(int n, const int* __restrict__ rp, const int* __restrict__ ci,
    const S* __restrict__ a, const double* __restrict__ dinv, const S* __restrict__ d, const S* __restrict__ xin,
    S* __restrict__ xout, double w, const int* __restrict__ run)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (run && !*run)) return;
    S rhs[3];
#pragma unroll
    for (int e = 0; e < 3; ++e) rhs[e] = d[3L * i + e];
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
        const int c = ci[k];
        if (c == i) continue;
        const S* blk = a + 9L * k;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            S s = blk[3 * e] * xin[3L * c];
            s += blk[3 * e + 1] * xin[3L * c + 1];
            s += blk[3 * e + 2] * xin[3L * c + 2];
            rhs[e] += S(-1) * s;
        }
    }
    const double* D = dinv + 9L * i;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        double o = D[3 * e] * (double)rhs[0];
        o += D[3 * e + 1] * (double)rhs[1];
        o += D[3 * e + 2] * (double)rhs[2];
        xout[3L * i + e] = (S)((1.0 - w) * (double)xin[3L * i + e] + w * (0.0 + 1.0 * o));
    }
}
template <class S>
void launchJacobiSweeps(int n, int sweeps, const int* rp, const int* ci, const S* a, const double* dinv, const S* d,
                        S* v, S* tmp, double w, const int* run, hipStream_t st)
{
    for (int k = 0; k < sweeps; ++k) {
        OM_HIP(hipMemcpyAsync(tmp, v, 3L * n * sizeof(S), hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL((k_jacobi_sweep<S>), dim3((n + 255) / 256), dim3(256), 0, st, n, rp, ci, a, dinv, d, tmp,
                           v, w, run);
    }
}
template void launchJacobiSweeps<double>(int, int, const int*, const int*, const double*, const double*,
                                         const double*, double*, double*, double, const int*, hipStream_t);
template void launchJacobiSweeps<float>(int, int, const int*, const int*, const float*, const double*, const float*,
                                        float*, float*, double, const int*, hipStream_t);

// ---------------------------------------------------------------------------
// Reductions: kRedBlocks fixed partials of 256 threads, then one block.
constexpr int kRedThreads = 256;

__device__ __forceinline__ double waveSum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ __forceinline__ double waveMax(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    return v;
}

template <bool IsMax>
__device__ __forceinline__ double blockReduce(double v)
{
    __shared__ double sm[kRedThreads / 64];
    v = IsMax ? waveMax(v) : waveSum(v);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) sm[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        r = sm[0];
        for (int q = 1; q < kRedThreads / 64; ++q) r = IsMax ? fmax(r, sm[q]) : r + sm[q];
    }
    __syncthreads();
    return r;
}

// final reduce of kRedBlocks partials with one block of kRedThreads
template <bool IsMax>
__device__ __forceinline__ double finalReduce(const double* __restrict__ partial)
{
    double v = IsMax ? 0.0 : 0.0;
    for (int k = threadIdx.x; k < kRedBlocks; k += kRedThreads) v = IsMax ? fmax(v, partial[k]) : v + partial[k];
    return blockReduce<IsMax>(v);
}

// Scalar products: the exact sum of the products, rounded once (exact_sum.h),
// so the value is the same for every launch shape and equals the CPU
// restatement's.  Every workgroup adds its products into limbs in LDS and
// those into the device accumulator; the last workgroup to finish rounds and
// evaluates the BiCGStab scalars of the product (one launch per product).
constexpr int kDotUnroll = 8;

inline int dotBlocks(long m)
{
    const long per = (long)kDotUnroll * kDotThreads;
    return (int)std::max(1L, std::min<long>(kDotMaxBlocks, (m + per - 1) / per));
}

__device__ __forceinline__ void markFail(KrylovScalars* ks, int tag)
{
    if (!ks->fail_code) { ks->fail_code = 2; ks->fail_tag = tag; }
}

// what the last workgroup does with the sums (s0, s1)
__device__ __forceinline__ void dotEpilogue(int mode, double s0, double s1, KrylovScalars* ks, double eps, int tagOff,
                                            double* out)
{
    if (mode == kDotValue) { *out = s0; return; }
    const int tag = ks->tag_base + tagOff;
    if (mode == kDotRho) {             // rho_i = <r0hat, r>, beta (bicgstabsolver.hh:197-206)
        ks->rho_i = s0;
        if (fabs(ks->rho) <= eps || fabs(ks->omega) <= eps) markFail(ks, tag);
        ks->beta = (s0 / ks->rho) * (ks->alpha / ks->omega);
        ks->rho = s0;
    } else if (mode == kDotAlpha) {    // denom = <r0hat, v>, alpha (:230-236)
        ks->denom = s0;
        if (fabs(s0) <= eps) markFail(ks, tag);
        ks->alpha = ks->rho / s0;
        if (fabs(ks->alpha) <= eps) markFail(ks, tag);
    } else {                           // omega = <t, s> / <t, t> (:287-292)
        if (fabs(s0) <= eps) markFail(ks, tag);
        ks->omega = s1 / s0;
        if (fabs(ks->omega) <= eps) markFail(ks, tag);
    }
}

// NACC = 1: sum a.b;  NACC = 2: sums a.b and a.c
template <class S, int NACC>
__global__ __launch_bounds__(kDotThreads) void k_xdot(long m, const S* __restrict__ a, const S* __restrict__ b,
    const S* __restrict__ c, DotScratch* __restrict__ sc, KrylovScalars* ks, int mode, double eps, int tagOff,
    double* out)
{
    if (ks && !ks->run) return;
    constexpr int nw = kDotThreads / 64;
    __shared__ unsigned long long sAcc[NACC * kAccLimbs];
    __shared__ unsigned sMeta[5];
    clearLimbs(sAcc, NACC, sMeta);
    __syncthreads();
    const int lane = (int)(threadIdx.x & 63u);
    const long wavesTotal = (long)gridDim.x * nw;
    unsigned fl0 = 0u, fl1 = 0u;
    // each wave takes kDotUnroll * 64 consecutive entries per pass (a uniform loop)
    for (long w = (long)blockIdx.x * nw + (long)(threadIdx.x >> 6); w * (kDotUnroll * 64) < m; w += wavesTotal) {
        const long base = w * (kDotUnroll * 64) + lane;
        double v0[kDotUnroll], v1[kDotUnroll];
#pragma unroll
        for (int u = 0; u < kDotUnroll; ++u) {
            const long i = base + 64L * u;
            v0[u] = 0.0; v1[u] = 0.0;
            if (i < m) {
                const double ai = (double)a[i];
                v0[u] = ai * (double)b[i];
                if (NACC == 2) v1[u] = ai * (double)c[i];
            }
        }
        accumulateBands<kDotUnroll>(v0, sAcc, fl0);
        if (NACC == 2) accumulateBands<kDotUnroll>(v1, sAcc + kAccLimbs, fl1);
    }
    if (fl0 | fl1) atomicOr(&sMeta[0], fl0 | (fl1 << 3));
    __syncthreads();
    flushLimbs(sAcc, NACC, sMeta, sc->acc);
    if (!lastToFinish(&sc->ticket, (int)gridDim.x, sMeta)) return;
    __shared__ long long sD[kAccLimbs];
    const double s0 = finishExact(&sc->acc[0], sD, sMeta);
    const double s1 = NACC == 2 ? finishExact(&sc->acc[1], sD, sMeta) : 0.0;
    if (threadIdx.x == 0) dotEpilogue(mode, s0, s1, ks, eps, tagOff, out);
}

template <class S>
void launchDotFinal(long m, const S* x, const S* y, DotScratch* sc, KrylovScalars* ks, int op, int tagOff,
                    hipStream_t st)
{
    const double eps = (double)std::numeric_limits<S>::min() * 1e10;
    hipLaunchKernelGGL((k_xdot<S, 1>), dim3(dotBlocks(m)), dim3(kDotThreads), 0, st, m, x, y, y, sc, ks,
                       op == 0 ? (int)kDotRho : (int)kDotAlpha, eps, tagOff, (double*)nullptr);
}
template void launchDotFinal<double>(long, const double*, const double*, DotScratch*, KrylovScalars*, int, int, hipStream_t);
template void launchDotFinal<float>(long, const float*, const float*, DotScratch*, KrylovScalars*, int, int, hipStream_t);

void launchScalarProduct(long m, const double* x, const double* y, DotScratch* sc, double* out, hipStream_t st)
{
    hipLaunchKernelGGL((k_xdot<double, 1>), dim3(dotBlocks(m)), dim3(kDotThreads), 0, st, m, x, y, y, sc,
                       (KrylovScalars*)nullptr, (int)kDotValue, 0.0, 0, out);
}

// p = r + (-beta) * (v*omega - p)   (bicgstabsolver.hh:211-222)
template <class S>
__global__ void k_update_p(long m, const S* __restrict__ r, const S* __restrict__ v, S* __restrict__ p,
                           const KrylovScalars* __restrict__ ks)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || !ks->run) return;
    const S omega = (S)ks->omega, beta = (S)ks->beta;
    S tmp = v[i];
    tmp *= omega;
    tmp -= p[i];
    tmp *= -beta;
    S pn = r[i];
    pn += tmp;
    p[i] = pn;
}
template <class S>
void launchUpdateP(long m, const S* r, const S* v, S* p, const KrylovScalars* ks, hipStream_t st)
{
    hipLaunchKernelGGL((k_update_p<S>), dim3((m + 255) / 256), dim3(256), 0, st, m, r, v, p, ks);
}
template void launchUpdateP<double>(long, const double*, const double*, double*, const KrylovScalars*, hipStream_t);
template void launchUpdateP<float>(long, const float*, const float*, float*, const KrylovScalars*, hipStream_t);

// x = x + alpha*y; r -= alpha*v; partials: max|r|, count(y != 0)  (:240-253)
template <class S>
__global__ __launch_bounds__(kRedThreads) void k_update_xr(long m, const S* __restrict__ y, const S* __restrict__ v,
    S* __restrict__ x, S* __restrict__ r, double* __restrict__ part, const KrylovScalars* __restrict__ ks)
{
    if (!ks->run) return;
    const S alpha = (S)ks->alpha;
    double mx = 0.0, nz = 0.0;
    for (long i = (long)blockIdx.x * kRedThreads + threadIdx.x; i < m; i += (long)kRedBlocks * kRedThreads) {
        const S yi = y[i];
        S tmp = yi;
        tmp *= alpha;
        tmp += x[i];
        x[i] = tmp;
        tmp = v[i];
        tmp *= alpha;
        S ri = r[i];
        ri -= tmp;
        r[i] = ri;
        mx = fmax(mx, fabs((double)ri));
        if (yi != S(0)) nz = 1.0;
    }
    mx = blockReduce<true>(mx);
    nz = blockReduce<true>(nz);
    if (threadIdx.x == 0) { part[blockIdx.x] = mx; part[kRedBlocks + blockIdx.x] = nz; }
}

__global__ __launch_bounds__(kRedThreads) void k_crit_final(const double* __restrict__ part, KrylovScalars* ks)
{
    if (!ks->run) return;
    const double mx = finalReduce<true>(part);
    const double nz = finalReduce<true>(part + kRedBlocks);
    if (threadIdx.x != 0) return;
    ks->resnorm = mx;
    ks->stag_half = (nz == 0.0) ? 1 : 0;
    // the test after the first half of iteration tag_base / 2 (bicgstabsolver.hh:
    // 252-273) with the outcome of the previous second half (:286-318), in the
    // order of the host loop; once it ends the solve, the queued kernels of
    // the next iteration return at once
    const int tag = ks->tag_base;
    int rc = -1, back = 0;
    if (tag > 0 && ks->fail_code && ks->fail_tag == tag - 1) { rc = 2; back = 1; }
    else if (tag > 0 && ks->stag_full_tag == tag - 1) { rc = 1; back = 1; }
    else if (ks->fail_code) rc = 2;
    else if (mx / fmax(1e-20, ks->init_norm) <= ks->tol || mx <= ks->abs_tol) rc = 0;
    else if (ks->stag_half || mx > ks->max_error) rc = 1;
    if (rc >= 0) {
        ks->stop_rc = rc;
        ks->stop_back = back;
        ks->run = 0;
    }
}

template <class S>
void launchUpdateXR(long m, const S* y, const S* v, S* x, S* r, double* partial, KrylovScalars* ks, hipStream_t st)
{
    hipLaunchKernelGGL((k_update_xr<S>), dim3(kRedBlocks), dim3(kRedThreads), 0, st, m, y, v, x, r, partial, ks);
    hipLaunchKernelGGL(k_crit_final, dim3(1), dim3(kRedThreads), 0, st, partial, ks);
}
template void launchUpdateXR<double>(long, const double*, const double*, double*, double*, double*, KrylovScalars*, hipStream_t);
template void launchUpdateXR<float>(long, const float*, const float*, float*, float*, double*, KrylovScalars*, hipStream_t);

// <t,t> and <t,s> in one pass; omega = <t,s>/<t,t>  (:287-292)
template <class S>
void launchDot2Final(long m, const S* t, const S* s, DotScratch* sc, KrylovScalars* ks, int tagOff, hipStream_t st)
{
    const double eps = (double)std::numeric_limits<S>::min() * 1e10;
    hipLaunchKernelGGL((k_xdot<S, 2>), dim3(dotBlocks(m)), dim3(kDotThreads), 0, st, m, t, t, s, sc, ks,
                       (int)kDotOmega, eps, tagOff, (double*)nullptr);
}
template void launchDot2Final<double>(long, const double*, const double*, DotScratch*, KrylovScalars*, int, hipStream_t);
template void launchDot2Final<float>(long, const float*, const float*, DotScratch*, KrylovScalars*, int, hipStream_t);

// x += omega*z (criterion change indicator z), r += -omega*t  (:296-325)
template <class S>
__global__ __launch_bounds__(kRedThreads) void k_update_xr2(long m, const S* __restrict__ z, const S* __restrict__ t,
    S* __restrict__ x, S* __restrict__ r, double* __restrict__ part, const KrylovScalars* __restrict__ ks)
{
    if (!ks->run) return;
    const S omega = (S)ks->omega;
    double nz = 0.0;
    for (long i = (long)blockIdx.x * kRedThreads + threadIdx.x; i < m; i += (long)kRedBlocks * kRedThreads) {
        const S zi = z[i];
        x[i] += omega * zi;
        if (zi != S(0)) nz = 1.0;
        r[i] += -omega * t[i];
    }
    nz = blockReduce<true>(nz);
    if (threadIdx.x == 0) part[blockIdx.x] = nz;
}
__global__ __launch_bounds__(kRedThreads) void k_stag_final(const double* __restrict__ part, KrylovScalars* ks,
                                                            int tagOff, int advance)
{
    if (!ks->run) return;
    const double nz = finalReduce<true>(part);
    if (threadIdx.x != 0) return;
    if (nz == 0.0 && ks->stag_full_tag < 0) ks->stag_full_tag = ks->tag_base + tagOff;
    ks->tag_base += advance;
}
template <class S>
void launchUpdateXR2(long m, const S* z, const S* t, S* x, S* r, double* partial, KrylovScalars* ks, int tagOff,
                     int advance, hipStream_t st)
{
    hipLaunchKernelGGL((k_update_xr2<S>), dim3(kRedBlocks), dim3(kRedThreads), 0, st, m, z, t, x, r, partial, ks);
    hipLaunchKernelGGL(k_stag_final, dim3(1), dim3(kRedThreads), 0, st, partial, ks, tagOff, advance);
}
template void launchUpdateXR2<double>(long, const double*, const double*, double*, double*, double*, KrylovScalars*, int, int, hipStream_t);
template void launchUpdateXR2<float>(long, const float*, const float*, float*, float*, double*, KrylovScalars*, int, int, hipStream_t);

// max |x| -> out (device)
template <class S>
__global__ __launch_bounds__(kRedThreads) void k_maxabs_partial(long m, const S* __restrict__ x, double* __restrict__ part)
{
    double mx = 0.0;
    for (long i = (long)blockIdx.x * kRedThreads + threadIdx.x; i < m; i += (long)kRedBlocks * kRedThreads)
        mx = fmax(mx, fabs((double)x[i]));
    mx = blockReduce<true>(mx);
    if (threadIdx.x == 0) part[blockIdx.x] = mx;
}
__global__ __launch_bounds__(kRedThreads) void k_max_final(const double* __restrict__ part, double* out)
{
    const double mx = finalReduce<true>(part);
    if (threadIdx.x == 0) *out = mx;
}
template <class S>
void launchMaxAbs(long m, const S* x, double* partial, double* out, hipStream_t st)
{
    hipLaunchKernelGGL((k_maxabs_partial<S>), dim3(kRedBlocks), dim3(kRedThreads), 0, st, m, x, partial);
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(kRedThreads), 0, st, partial, out);
}
template void launchMaxAbs<double>(long, const double*, double*, double*, hipStream_t);
template void launchMaxAbs<float>(long, const float*, double*, double*, hipStream_t);

// r = b - A x  (FP64 outer residual of the mixed-precision refinement)
template <int BS>
__global__ __launch_bounds__(256) void k_residual(int n, const int* __restrict__ rp, const int* __restrict__ ci,
    const double* __restrict__ a, const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ r)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc[BS];
#pragma unroll
    for (int e = 0; e < BS; ++e) acc[e] = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
        const int c = ci[k];
        const double* blk = a + (long)BS * BS * k;
#pragma unroll
        for (int e = 0; e < BS; ++e) {
            double s = blk[e * BS] * x[(long)BS * c];
#pragma unroll
            for (int q = 1; q < BS; ++q) s += blk[e * BS + q] * x[(long)BS * c + q];
            acc[e] += s;
        }
    }
#pragma unroll
    for (int e = 0; e < BS; ++e) r[(long)BS * i + e] = b[(long)BS * i + e] - acc[e];
}
void launchResidualUpdate(int n, int bs, const int* rp, const int* ci, const double* a, const double* x,
                          const double* b, double* r, hipStream_t st)
{
    dim3 g((n + 255) / 256), blk(256);
    if (bs == 3) hipLaunchKernelGGL(k_residual<3>, g, blk, 0, st, n, rp, ci, a, x, b, r);
    else hipLaunchKernelGGL(k_residual<2>, g, blk, 0, st, n, rp, ci, a, x, b, r);
}

__global__ void k_axpy(long m, double alpha, const double* __restrict__ x, double* __restrict__ y)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) y[i] += alpha * x[i];
}
void launchAxpyD(long m, double alpha, const double* x, double* y, hipStream_t st)
{
    hipLaunchKernelGGL(k_axpy, dim3((m + 255) / 256), dim3(256), 0, st, m, alpha, x, y);
}
__global__ void k_add_float(long m, const float* __restrict__ d, double* __restrict__ x)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) x[i] += (double)d[i];
}
void launchAddFromFloat(long m, const float* d, double* x, hipStream_t st)
{
    hipLaunchKernelGGL(k_add_float, dim3((m + 255) / 256), dim3(256), 0, st, m, d, x);
}

}  // namespace omodels
