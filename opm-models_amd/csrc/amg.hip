// amg.hip -- aggregation AMG of a 3x3 system on the 7-point grid, the
// preconditioner of precond = 2.
//
// The reference's ParallelAmgBackend (parallelamgbackend.hh:85-316) runs its
// BiCGStab with dune-istl's aggregation AMG: SOR smoother with one iteration
// and relaxation 1 (:272-276), aggregates of about three cells per direction
// (:294-295), coarsening until AmgCoarsenTarget cells (:72, :290), a direct
// solver on the coarsest level.  Here the grid itself gives the aggregates:
//   - box aggregates of ax x ay x az cells; every level is again a 7-point
//     grid, so one set of kernels serves every level;
//   - Galerkin coarse operators P^T A P with piecewise-constant P: a coarse
//     block sums the fine blocks between the two aggregates, over the
//     aggregate's cells in natural order and each cell's slots in order;
//   - a multiplicative V-cycle: `pre` forward block Gauss-Seidel sweeps from
//     x = 0, the residual summed over each aggregate as the coarse right-hand
//     side, the coarse solution prolongated with damping, `post` backward
//     sweeps.  The residual is formed and restricted on the even cells only:
//     the odd ones, swept last, solve their rows with their neighbours' final
//     values, so theirs is zero but for rounding;
//   - Gauss-Seidel in red-black order: the cells with i + j + k even, then
//     the odd ones (backward: odd first); a cell's six neighbours have the
//     other colour, so a colour is one launch with one independent lane per
//     cell.  Each level keeps its matrix in a red-black layout [7][N][9]
//     whose rows are the even cells in natural order, then the odd ones:
//     the lanes of a colour read consecutive blocks of every slot (the
//     natural layout would give them every other block).  Vectors stay in
//     natural order;
//   - the coarsest level (at most kAmgMaxCoarse cells) solved by its dense
//     inverse, computed in place by Gauss-Jordan with partial pivoting in one
//     workgroup's LDS.
// Every kernel fixes its operation order, which the CPU restatement follows.
#include "amg.h"

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "ilu_colored.h"

namespace omodels {

namespace {

struct AmgGrid {
    int nx, ny, nz;
    long N;
};

AmgGrid gridOf(int nx, int ny, int nz) { return {nx, ny, nz, (long)nx * ny * nz}; }

// neighbour of cell I = (i, j, k) through slot s, or -1
__device__ __forceinline__ long slotNeighbour(const AmgGrid& g, long I, int i, int j, int k, int s)
{
    const long nxy = (long)g.nx * g.ny;
    switch (s) {
    case 0: return k > 0 ? I - nxy : -1;
    case 1: return j > 0 ? I - g.nx : -1;
    case 2: return i > 0 ? I - 1 : -1;
    case 3: return I;
    case 4: return i + 1 < g.nx ? I + 1 : -1;
    case 5: return j + 1 < g.ny ? I + g.nx : -1;
    default: return k + 1 < g.nz ? I + nxy : -1;
    }
}

// r = b - A x on the rows [0, N0) of the red-black layout (the even cells);
// one lane per row (red-black row L = t / 3 of cell I, row e = t % 3), slots
// in order
template <class S>
__global__ void k_amg_residual(AmgGrid g, long nRows, const int* __restrict__ rbInv, const S* __restrict__ A,
                               const S* __restrict__ b, const S* __restrict__ x, S* __restrict__ r,
                               const int* __restrict__ run)
{
    if (run && !*run) return;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * nRows) return;
    const long L = t / 3;
    const int e = (int)(t - 3 * L);
    const long I = rbInv[L];
    const int i = (int)(I % g.nx), j = (int)((I / g.nx) % g.ny), k = (int)(I / ((long)g.nx * g.ny));
    S acc = b[3 * I + e];
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        const long J = slotNeighbour(g, I, i, j, k, s);
        if (J < 0) continue;
        const S* a = A + 9 * (s * g.N + L) + 3 * e;
        S v = a[0] * x[3 * J];
        v += a[1] * x[3 * J + 1];
        v += a[2] * x[3 * J + 2];
        acc += S(-1) * v;
    }
    r[3 * I + e] = acc;
}

// One colour of a block Gauss-Seidel sweep, in place: the rows [base,
// base + count) of the red-black layout, x_I = A_II^-1 (b_I - sum_{K != I}
// A_IK x_K), the neighbours in column order; skip: the neighbours' values
// are zero (the first colour of a sweep from x = 0) and are not read.  The
// inverted diagonal blocks are FP64 in both precisions.
template <class S>
__global__ void k_amg_rb_gs(AmgGrid g, long base, long count, const int* __restrict__ rbInv,
                            const S* __restrict__ A, const double* __restrict__ dinv,
                            const S* __restrict__ b, S* __restrict__ x, int skip,
                            const int* __restrict__ run)
{
    if (run && !*run) return;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const long L = base + t;
    const long I = rbInv[L];
    const int i = (int)(I % g.nx), j = (int)((I / g.nx) % g.ny), k = (int)(I / ((long)g.nx * g.ny));
    S acc[3] = {b[3 * I], b[3 * I + 1], b[3 * I + 2]};
    if (!skip) {
        constexpr int kOrder[6] = {0, 1, 2, 4, 5, 6};
#pragma unroll
        for (int o = 0; o < 6; ++o) {
            const int s = kOrder[o];
            const long K = slotNeighbour(g, I, i, j, k, s);
            if (K < 0) continue;
            const S* a = A + 9 * (s * g.N + L);
            const S x0 = x[3 * K], x1 = x[3 * K + 1], x2 = x[3 * K + 2];
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                S v = a[3 * e] * x0;
                v += a[3 * e + 1] * x1;
                v += a[3 * e + 2] * x2;
                acc[e] += S(-1) * v;
            }
        }
    }
    const double* Di = dinv + 9 * L;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        double o = Di[3 * e] * (double)acc[0];
        o += Di[3 * e + 1] * (double)acc[1];
        o += Di[3 * e + 2] * (double)acc[2];
        x[3 * I + e] = (S)(0.0 + o);
    }
}

// the natural structured layout in the red-black one: one lane per (row, entry)
__global__ void k_amg_to_rb(AmgGrid g, const int* __restrict__ rbInv, const double* __restrict__ A,
                            double* __restrict__ Arb)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 63 * g.N) return;
    const long s = t / (9 * g.N);
    const long rem = t - s * 9 * g.N;
    const long L = rem / 9;
    const int q = (int)(rem - 9 * L);
    Arb[t] = A[9 * (s * g.N + rbInv[L]) + q];
}

// coarse right-hand side: the sum of r over each aggregate, natural order
template <class S>
__global__ void k_amg_restrict(AmgGrid f, AmgGrid c, int ax, int ay, int az, const S* __restrict__ r,
                               S* __restrict__ bc, const int* __restrict__ run)
{
    if (run && !*run) return;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * c.N) return;
    const long a = t / 3;
    const int e = (int)(t - 3 * a);
    const int ci = (int)(a % c.nx), cj = (int)((a / c.nx) % c.ny), ck = (int)(a / ((long)c.nx * c.ny));
    S acc = S(0);
    for (int k = ck * az; k < min(f.nz, ck * az + az); ++k)
        for (int j = cj * ay; j < min(f.ny, cj * ay + ay); ++j)
            for (int i = ci * ax; i < min(f.nx, ci * ax + ax); ++i)
                if (((i + j + k) & 1) == 0) acc += r[3 * (i + (long)f.nx * (j + (long)f.ny * k)) + e];
    bc[t] = acc;
}

// x += damp * P xc
template <class S>
__global__ void k_amg_prolong(AmgGrid f, AmgGrid c, int ax, int ay, int az, double damp,
                              const S* __restrict__ xc, S* __restrict__ x, const int* __restrict__ run)
{
    if (run && !*run) return;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * f.N) return;
    const long I = t / 3;
    const int e = (int)(t - 3 * I);
    const int i = (int)(I % f.nx), j = (int)((I / f.nx) % f.ny), k = (int)(I / ((long)f.nx * f.ny));
    const long a = i / ax + (long)c.nx * (j / ay + (long)c.ny * (k / az));
    x[t] = x[t] + (S)damp * xc[3 * a + e];
}

// coarse operator: one lane per (coarse slot t, coarse cell, block entry).
// The fine blocks of the aggregate's cells (natural order) and slots (in
// order) are summed into the coarse slot of the neighbour's aggregate; each
// lane forms one of those sums, so it visits only the cells that feed it:
// for a face slot the aggregate's layer on that face, for the diagonal all
// cells and their slots whose neighbour stays inside.  rbOfF: nullptr for a
// fine matrix in the natural layout, else its red-black rows.
__global__ void k_amg_galerkin(AmgGrid f, AmgGrid c, int ax, int ay, int az, const int* __restrict__ rbOfF,
                               const int* __restrict__ rbOfC, const double* __restrict__ Af,
                               double* __restrict__ Ac)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= 63 * c.N) return;
    const int t = (int)(g / (9 * c.N));
    const long rem = g - (long)t * 9 * c.N;
    const long a = rem / 9;
    const int q = (int)(rem - 9 * a);
    const int ci = (int)(a % c.nx), cj = (int)((a / c.nx) % c.ny), ck = (int)(a / ((long)c.nx * c.ny));
    int k0 = ck * az, k1 = min(f.nz, ck * az + az), j0 = cj * ay, j1 = min(f.ny, cj * ay + ay);
    int i0 = ci * ax, i1 = min(f.nx, ci * ax + ax);
    // a face slot: only the aggregate's layer on that face (absent when the
    // aggregate is clipped by the grid's edge on that side)
    switch (t) {
    case 0: k1 = k0 + 1; break;
    case 1: j1 = j0 + 1; break;
    case 2: i1 = i0 + 1; break;
    case 4: i0 = ci * ax + ax - 1; break;
    case 5: j0 = cj * ay + ay - 1; break;
    case 6: k0 = ck * az + az - 1; break;
    default: break;
    }
    double acc = 0.0;
    for (int k = k0; k < k1; ++k)
        for (int j = j0; j < j1; ++j)
            for (int i = i0; i < i1; ++i) {
                const long I = i + (long)f.nx * (j + (long)f.ny * k);
                const long row = rbOfF ? rbOfF[I] : I;
                if (t != 3) {
                    if (slotNeighbour(f, I, i, j, k, t) >= 0) acc += Af[9 * (t * f.N + row) + q];
                    continue;
                }
#pragma unroll
                for (int s = 0; s < 7; ++s) {
                    if (slotNeighbour(f, I, i, j, k, s) < 0) continue;
                    // the neighbour lies in this aggregate unless the slot
                    // crosses the aggregate's face in its direction
                    bool same;
                    switch (s) {
                    case 0: same = k > ck * az; break;
                    case 1: same = j > cj * ay; break;
                    case 2: same = i > ci * ax; break;
                    case 3: same = true; break;
                    case 4: same = i + 1 < ci * ax + ax; break;
                    case 5: same = j + 1 < cj * ay + ay; break;
                    default: same = k + 1 < ck * az + az; break;
                    }
                    if (same) acc += Af[9 * (s * f.N + row) + q];
                }
            }
    Ac[9 * (t * c.N + rbOfC[a]) + q] = acc;
}

// the coarsest level as a dense row-major n x n matrix, n = 3N
__global__ void k_amg_dense_build(AmgGrid g, const int* __restrict__ rbOf, const double* __restrict__ A,
                                  double* __restrict__ M)
{
    const long n = 3 * g.N;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * n) return;
    const long row = t / n, col = t - row * n;
    const long I = row / 3, J = col / 3;
    const int e = (int)(row - 3 * I), f = (int)(col - 3 * J);
    const int i = (int)(I % g.nx), j = (int)((I / g.nx) % g.ny), k = (int)(I / ((long)g.nx * g.ny));
    double v = 0.0;
    for (int s = 0; s < 7; ++s)
        if (slotNeighbour(g, I, i, j, k, s) == J) { v = A[9 * (s * g.N + rbOf[I]) + 3 * e + f]; break; }
    M[t] = v;
}

constexpr int kGjThreads = 1024;
constexpr int kDenseMax = 3 * kAmgMaxCoarse;
constexpr int kDenseLd = kDenseMax + 1;      // LDS row stride: column reads hit distinct banks

// Gauss-Jordan inverse of the dense n x n matrix M with partial pivoting (the
// first row of the largest |entry| of the column), in one workgroup's LDS;
// the inverse is written back transposed (M[j n + i] = M^-1_ij, so that the
// application reads it by columns); fail |= 2 when an entry is not finite or
// a pivot is zero.  Step k on the matrix S after steps 0..k-1: rows k and p
// exchanged, row k scaled by r = 1 / S_pk,
//     S'_kj = (j == k ? 1 : S_pj) r,
//     S'_ij = (j == k ? 0 : S_sigma(i)j) - S'_kj S_sigma(i)k   (i != k),
// sigma exchanging k and p; a step reads one LDS copy and writes the other.
// The column exchanges of the pivots are applied when the result is written.
__global__ __launch_bounds__(kGjThreads) void k_amg_dense_invert(int n, double* __restrict__ M, int* __restrict__ fail)
{
    __shared__ double S[2][kDenseMax * kDenseLd];
    __shared__ int piv[kDenseMax];
    __shared__ int col[kDenseMax];
    __shared__ int bad;
    const int tid = (int)threadIdx.x;
    const int nn = n * n;
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int q = tid; q < nn; q += kGjThreads) {
        const int i = q / n, j = q - i * n;
        S[0][i * kDenseLd + j] = M[q];
        if (!isfinite(M[q])) bad = 1;
    }
    __syncthreads();
    if (bad) {
        if (tid == 0) atomicOr(fail, 2);
        return;
    }
    // thread (rg, j) updates column j of rows rg, rg + ngroups, ...: the
    // index arithmetic once, not per element
    const int jc = tid % n, rg = tid / n, ngroups = kGjThreads / n;
    int cur = 0;
    for (int k = 0; k < n; ++k) {
        const double* A = S[cur];
        double* B = S[cur ^ 1];
        if (tid < 64) {
            // pivot search by one wave: lane l looks at rows k + l, k + l + 64
            double best = -1.0;
            int bi = n;
            for (int i = k + tid; i < n; i += 64) {
                const double v = fabs(A[i * kDenseLd + k]);
                if (v > best || (v == best && i < bi)) { best = v; bi = i; }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_down(best, off, 64);
                const int oi = __shfl_down(bi, off, 64);
                if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
            }
            if (tid == 0) {
                piv[k] = bi;
                if (!(best > 0.0)) bad = 1;
            }
        }
        __syncthreads();
        if (bad) {
            if (tid == 0) atomicOr(fail, 2);
            return;
        }
        const int p = piv[k];
        const double r = 1.0 / A[p * kDenseLd + k];
        if (rg < ngroups) {
            const int j = jc;
            const double rowk = (j == k ? 1.0 : A[p * kDenseLd + j]) * r;
            for (int i = rg; i < n; i += ngroups) {
                if (i == k) {
                    B[i * kDenseLd + j] = rowk;
                } else {
                    const int si = i == p ? k : i;
                    B[i * kDenseLd + j] = (j == k ? 0.0 : A[si * kDenseLd + j]) - rowk * A[si * kDenseLd + k];
                }
            }
        }
        cur ^= 1;
        __syncthreads();
    }
    // the column exchanges k <-> piv[k], k descending: result column c holds column col[c]
    if (tid == 0) {
        for (int c = 0; c < n; ++c) col[c] = c;
        for (int k = n - 1; k >= 0; --k) {
            const int t = col[k];
            col[k] = col[piv[k]];
            col[piv[k]] = t;
        }
    }
    __syncthreads();
    for (int q = tid; q < nn; q += kGjThreads) {
        const int i = q / n, c = q - i * n;
        M[c * n + i] = S[cur][i * kDenseLd + col[c]];
    }
}

// x = M^-1 b on the coarsest level (the inverse stored transposed): the
// inverse and b staged in LDS, then one lane per row, columns in order
template <class S>
__global__ __launch_bounds__(kGjThreads) void k_amg_dense_apply(int n, const double* __restrict__ MinvT,
                                                                const S* __restrict__ b, S* __restrict__ x,
                                                                const int* __restrict__ run)
{
    if (run && !*run) return;
    __shared__ double M[kDenseMax * kDenseMax];
    __shared__ double bs[kDenseMax];
    const int tid = (int)threadIdx.x;
    for (int q = tid; q < n * n; q += kGjThreads) M[q] = MinvT[q];
    if (tid < n) bs[tid] = (double)b[tid];
    __syncthreads();
    if (tid >= n) return;
    double s = M[tid] * bs[0];
    for (int j = 1; j < n; ++j) s += M[j * n + tid] * bs[j];
    x[tid] = (S)s;
}

// FP32 copy of a level matrix
__global__ void k_amg_to_float(long m, const double* __restrict__ a, float* __restrict__ f)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) f[t] = (float)a[t];
}

unsigned blocksFor(long threads, int per) { return (unsigned)((threads + per - 1) / per); }

}  // namespace

void checkAmgParams(const AmgParams& p)
{
    if (p.coarsenTarget < 1 || p.coarsenTarget > kAmgMaxCoarse)
        throw std::invalid_argument("AMG: the coarsening target must be in [1, " + std::to_string(kAmgMaxCoarse) +
                                    "] cells (the coarsest level is solved by a dense inverse)");
    if (p.pre < 1 || p.post < 0) throw std::invalid_argument("AMG: pre-smoothing >= 1, post-smoothing >= 0 sweeps");
    if (p.agg[0] < 1 || p.agg[1] < 1 || p.agg[2] < 1) throw std::invalid_argument("AMG: aggregate sizes >= 1");
    if (!std::isfinite(p.damp)) throw std::invalid_argument("AMG: prolongation damping must be finite");
}

std::vector<std::array<int, 3>> amgLevelDims(int nx, int ny, int nz, const AmgParams& p)
{
    std::vector<std::array<int, 3>> d{{nx, ny, nz}};
    while ((long)d.back()[0] * d.back()[1] * d.back()[2] > p.coarsenTarget && d.size() < 16) {
        const auto f = d.back();
        const std::array<int, 3> c{(f[0] + p.agg[0] - 1) / p.agg[0], (f[1] + p.agg[1] - 1) / p.agg[1],
                                   (f[2] + p.agg[2] - 1) / p.agg[2]};
        if ((long)c[0] * c[1] * c[2] == (long)f[0] * f[1] * f[2]) break;
        d.push_back(c);
    }
    return d;
}

AmgHierarchy::AmgHierarchy(int nx, int ny, int nz, const AmgParams& p, bool fp32, hipStream_t s)
    : p_(p), fp32_(fp32)
{
    checkAmgParams(p);
    const auto dims = amgLevelDims(nx, ny, nz, p);
    if ((long)dims.back()[0] * dims.back()[1] * dims.back()[2] > kAmgMaxCoarse)
        throw std::invalid_argument("AMG: the grid does not coarsen to " + std::to_string(kAmgMaxCoarse) +
                                    " cells within 16 levels");
    for (size_t l = 0; l < dims.size(); ++l) {
        lv_.push_back(std::make_unique<Level>());
        Level& L = *lv_[l];
        L.nx = dims[l][0]; L.ny = dims[l][1]; L.nz = dims[l][2];
        L.N = (long)L.nx * L.ny * L.nz;
        if (L.N > INT32_MAX) throw std::invalid_argument("AMG: grid too large");
        const bool last = l + 1 == dims.size();
        // red-black rows: the even cells (i + j + k even) in natural order, then the odd ones
        std::vector<int> inv(L.N);
        L.hostRbOf.resize(L.N);
        long e = 0;
        for (int pass = 0; pass < 2; ++pass) {
            for (long I = 0; I < L.N; ++I) {
                const int i = (int)(I % L.nx), j = (int)((I / L.nx) % L.ny), k = (int)(I / ((long)L.nx * L.ny));
                if (((i + j + k) & 1) != pass) continue;
                inv[e] = (int)I;
                L.hostRbOf[I] = (int)e;
                ++e;
            }
            if (pass == 0) L.N0 = e;
        }
        L.rbInv.upload(inv.data(), inv.size(), s);
        L.rbOf.upload(L.hostRbOf.data(), L.hostRbOf.size(), s);
        L.A.alloc(63 * L.N);
        if (!last) {
            L.ax = p.agg[0]; L.ay = p.agg[1]; L.az = p.agg[2];
            L.dinv.alloc(9 * L.N);
            L.r.alloc(3 * L.N);
        }
        if (l > 0) {
            L.b.alloc(3 * L.N);
            L.x.alloc(3 * L.N);
        }
        if (fp32) {
            L.Af.alloc(63 * L.N);
            if (!last) L.rf.alloc(3 * L.N);
            if (l > 0) { L.bf.alloc(3 * L.N); L.xf.alloc(3 * L.N); }
        }
    }
    const long n = 3 * lv_.back()->N;
    dense_.alloc(n * n);
    OM_HIP(hipStreamSynchronize(s));     // the host copies of the maps are uploaded from
}

void AmgHierarchy::setup(const double* A0, int* fail, hipStream_t s)
{
    {
        Level& F = *lv_[0];
        const AmgGrid f = gridOf(F.nx, F.ny, F.nz);
        hipLaunchKernelGGL(k_amg_to_rb, dim3(blocksFor(63 * f.N, 256)), dim3(256), 0, s, f, F.rbInv.p, A0, F.A.p);
    }
    for (size_t l = 0; l < lv_.size(); ++l) {
        Level& L = *lv_[l];
        if (l > 0) {
            const Level& F = *lv_[l - 1];
            const AmgGrid f = gridOf(F.nx, F.ny, F.nz), c = gridOf(L.nx, L.ny, L.nz);
            // the first coarse level from the caller's natural layout (whole
            // cache lines), the others from the red-black copies
            hipLaunchKernelGGL(k_amg_galerkin, dim3(blocksFor(63 * c.N, 256)), dim3(256), 0, s, f, c, F.ax, F.ay,
                               F.az, l == 1 ? nullptr : F.rbOf.p, L.rbOf.p, l == 1 ? A0 : F.A.p, L.A.p);
        }
        if (l + 1 < lv_.size()) launchDiaBlockInverse(L.N, L.A.p, L.dinv.p, fail, s);
    }
    const Level& C = *lv_.back();
    const AmgGrid c = gridOf(C.nx, C.ny, C.nz);
    const long n = 3 * C.N;
    hipLaunchKernelGGL(k_amg_dense_build, dim3(blocksFor(n * n, 256)), dim3(256), 0, s, c, C.rbOf.p, C.A.p,
                       dense_.p);
    hipLaunchKernelGGL(k_amg_dense_invert, dim3(1), dim3(kGjThreads), 0, s, (int)n, dense_.p, fail);
    if (fp32_)
        for (auto& L : lv_)
            hipLaunchKernelGGL(k_amg_to_float, dim3(blocksFor(63 * L->N, 256)), dim3(256), 0, s, 63 * L->N, L->A.p,
                               L->Af.p);
    OM_HIP(hipGetLastError());
    ready_ = true;
}

template <class S>
void AmgHierarchy::cycle(size_t l, const S* b, S* x, const int* run, hipStream_t s)
{
    Level& L = *lv_[l];
    if (l + 1 == lv_.size()) {
        const int n = (int)(3 * L.N);
        hipLaunchKernelGGL((k_amg_dense_apply<S>), dim3(1), dim3(kGjThreads), 0, s, n, dense_.p, b, x, run);
        return;
    }
    Level& C = *lv_[l + 1];
    const AmgGrid f = gridOf(L.nx, L.ny, L.nz), c = gridOf(C.nx, C.ny, C.nz);
    const S* A = L.mat<S>();
    S* r = L.res<S>();
    S* cb = C.rhs<S>();
    S* cx = C.sol<S>();
    // one colour: rows [base, base + count) of the red-black layout
    auto colour = [&](int col, bool skip) {
        const long base = col == 0 ? 0 : L.N0, count = col == 0 ? L.N0 : L.N - L.N0;
        if (count == 0) return;
        hipLaunchKernelGGL((k_amg_rb_gs<S>), dim3(blocksFor(count, 256)), dim3(256), 0, s, f, base, count,
                           L.rbInv.p, A, L.dinv.p, b, x, skip ? 1 : 0, run);
    };
    for (int k = 0; k < p_.pre; ++k) {
        colour(0, k == 0);          // from x = 0 the even cells' neighbours are still zero
        colour(1, false);
    }
    hipLaunchKernelGGL((k_amg_residual<S>), dim3(blocksFor(3 * L.N0, 256)), dim3(256), 0, s, f, L.N0, L.rbInv.p, A,
                       b, x, r, run);
    hipLaunchKernelGGL((k_amg_restrict<S>), dim3(blocksFor(3 * c.N, 256)), dim3(256), 0, s, f, c, L.ax, L.ay, L.az,
                       (const S*)r, cb, run);
    cycle<S>(l + 1, cb, cx, run, s);
    hipLaunchKernelGGL((k_amg_prolong<S>), dim3(blocksFor(3 * f.N, 256)), dim3(256), 0, s, f, c, L.ax, L.ay, L.az,
                       p_.damp, (const S*)cx, x, run);
    for (int k = 0; k < p_.post; ++k) {
        colour(1, false);
        colour(0, false);
    }
}

template <class S>
void AmgHierarchy::apply(const S* d, S* v, const int* run, hipStream_t s)
{
    if (!ready_) throw std::logic_error("AMG: apply before setup");
    if (sizeof(S) == 4 && !fp32_) throw std::logic_error("AMG: no FP32 copy of the hierarchy");
    cycle<S>(0, d, v, run, s);
    OM_HIP(hipGetLastError());
}
template void AmgHierarchy::apply<double>(const double*, double*, const int*, hipStream_t);
template void AmgHierarchy::apply<float>(const float*, float*, const int*, hipStream_t);

void AmgHierarchy::downloadLevel(int l, double* host, hipStream_t s) const
{
    const Level& L = *lv_[l];
    std::vector<double> rb(63 * L.N);
    OM_HIP(hipMemcpyAsync(rb.data(), L.A.p, rb.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    OM_HIP(hipStreamSynchronize(s));
    for (int sl = 0; sl < 7; ++sl)
        for (long I = 0; I < L.N; ++I)
            for (int q = 0; q < 9; ++q) host[9 * (sl * L.N + I) + q] = rb[9 * (sl * L.N + L.hostRbOf[I]) + q];
}

}  // namespace omodels
