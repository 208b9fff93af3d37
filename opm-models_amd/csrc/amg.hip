// amg.hip -- aggregation AMG of a 3x3 system on the 7-point grid, the
// preconditioner of precond = 2.
//
// The reference's ParallelAmgBackend (parallelamgbackend.hh:85-316) runs its
// BiCGStab with dune-istl's aggregation AMG: SOR smoother with one iteration
// and relaxation 1 (:272-276), aggregates of about three cells per direction
// (:294-295), coarsening until AmgCoarsenTarget cells (:72, :290), a direct
// solver on the coarsest level.  Here the grid itself gives the aggregates:
//   - box aggregates of ax x ay x az cells; every level is again a 7-point
//     grid, its matrix in the structured layout [7][N][9] (slots -z, -y, -x,
//     I, +x, +y, +z), so the fine level's kernels serve every level;
//   - Galerkin coarse operators P^T A P with piecewise-constant P: a coarse
//     block sums the fine blocks between the two aggregates, over the
//     aggregate's cells in natural order and each cell's slots in order;
//   - a multiplicative V-cycle: `pre` forward block Gauss-Seidel sweeps from
//     x = 0, the residual summed over each aggregate as the coarse right-hand
//     side, the coarse solution prolongated with damping, `post` backward
//     sweeps; Gauss-Seidel in the two-colour tile order of ilu_colored.hip
//     (every tile of a colour at once, the sweep a pass over the matrix);
//   - the coarsest level (at most kAmgMaxCoarse cells) solved by its dense
//     inverse, computed in place by Gauss-Jordan with partial pivoting in one
//     workgroup.
// Every kernel fixes its operation order, which the CPU restatement follows.
#include "amg.h"

#include <cmath>
#include <stdexcept>
#include <string>

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

// r = b - A x; one lane per row (cell t / 3, row t % 3), slots in order
__global__ void k_amg_residual(AmgGrid g, const double* __restrict__ A, const double* __restrict__ b,
                               const double* __restrict__ x, double* __restrict__ r, const int* __restrict__ run)
{
    if (run && !*run) return;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * g.N) return;
    const long I = t / 3;
    const int e = (int)(t - 3 * I);
    const int i = (int)(I % g.nx), j = (int)((I / g.nx) % g.ny), k = (int)(I / ((long)g.nx * g.ny));
    double acc = b[t];
#pragma unroll
    for (int s = 0; s < 7; ++s) {
        const long J = slotNeighbour(g, I, i, j, k, s);
        if (J < 0) continue;
        const double* a = A + 9 * (s * g.N + I) + 3 * e;
        double v = a[0] * x[3 * J];
        v += a[1] * x[3 * J + 1];
        v += a[2] * x[3 * J + 2];
        acc += -1.0 * v;
    }
    r[t] = acc;
}

// coarse right-hand side: the sum of r over each aggregate, natural order
__global__ void k_amg_restrict(AmgGrid f, AmgGrid c, int ax, int ay, int az, const double* __restrict__ r,
                               double* __restrict__ bc, const int* __restrict__ run)
{
    if (run && !*run) return;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * c.N) return;
    const long a = t / 3;
    const int e = (int)(t - 3 * a);
    const int ci = (int)(a % c.nx), cj = (int)((a / c.nx) % c.ny), ck = (int)(a / ((long)c.nx * c.ny));
    double acc = 0.0;
    for (int k = ck * az; k < min(f.nz, ck * az + az); ++k)
        for (int j = cj * ay; j < min(f.ny, cj * ay + ay); ++j)
            for (int i = ci * ax; i < min(f.nx, ci * ax + ax); ++i)
                acc += r[3 * (i + (long)f.nx * (j + (long)f.ny * k)) + e];
    bc[t] = acc;
}

// x += damp * P xc
__global__ void k_amg_prolong(AmgGrid f, AmgGrid c, int ax, int ay, int az, double damp,
                              const double* __restrict__ xc, double* __restrict__ x, const int* __restrict__ run)
{
    if (run && !*run) return;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * f.N) return;
    const long I = t / 3;
    const int e = (int)(t - 3 * I);
    const int i = (int)(I % f.nx), j = (int)((I / f.nx) % f.ny), k = (int)(I / ((long)f.nx * f.ny));
    const long a = i / ax + (long)c.nx * (j / ay + (long)c.ny * (k / az));
    x[t] = x[t] + damp * xc[3 * a + e];
}

// coarse operator: one lane per (coarse cell, block entry); the fine blocks
// of the aggregate's cells (natural order) and slots (in order) summed into
// the coarse slot of the neighbour's aggregate
__global__ void k_amg_galerkin(AmgGrid f, AmgGrid c, int ax, int ay, int az, const double* __restrict__ Af,
                               double* __restrict__ Ac)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 9 * c.N) return;
    const long a = t / 9;
    const int q = (int)(t - 9 * a);
    const int ci = (int)(a % c.nx), cj = (int)((a / c.nx) % c.ny), ck = (int)(a / ((long)c.nx * c.ny));
    double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = ck * az; k < min(f.nz, ck * az + az); ++k)
        for (int j = cj * ay; j < min(f.ny, cj * ay + ay); ++j)
            for (int i = ci * ax; i < min(f.nx, ci * ax + ax); ++i) {
                const long I = i + (long)f.nx * (j + (long)f.ny * k);
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
                    acc[same ? 3 : s] += Af[9 * (s * f.N + I) + q];
                }
            }
#pragma unroll
    for (int s = 0; s < 7; ++s) Ac[9 * (s * c.N + a) + q] = acc[s];
}

// the coarsest level as a dense row-major n x n matrix, n = 3N
__global__ void k_amg_dense_build(AmgGrid g, const double* __restrict__ A, double* __restrict__ M)
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
        if (slotNeighbour(g, I, i, j, k, s) == J) { v = A[9 * (s * g.N + I) + 3 * e + f]; break; }
    M[t] = v;
}

constexpr int kGjThreads = 256;

// in-place Gauss-Jordan inverse of the dense n x n matrix with partial
// pivoting (the first row of the largest |entry| of the column), one
// workgroup; fail |= 2 when an entry is not finite or a pivot is zero
__global__ __launch_bounds__(kGjThreads) void k_amg_dense_invert(int n, double* __restrict__ M, int* __restrict__ fail)
{
    __shared__ double fcol[3 * kAmgMaxCoarse];
    __shared__ int piv[3 * kAmgMaxCoarse];
    __shared__ double wv[kGjThreads / 64];
    __shared__ int wi[kGjThreads / 64];
    __shared__ int bad;
    const int tid = (int)threadIdx.x;
    const long nn = (long)n * n;
    if (tid == 0) bad = 0;
    __syncthreads();
    for (long q = tid; q < nn; q += kGjThreads)
        if (!isfinite(M[q])) bad = 1;
    __syncthreads();
    if (bad) {
        if (tid == 0) atomicOr(fail, 2);
        return;
    }
    for (int k = 0; k < n; ++k) {
        double best = -1.0;
        int bi = n;
        for (int i = k + tid; i < n; i += kGjThreads) {
            const double v = fabs(M[(long)i * n + k]);
            if (v > best || (v == best && i < bi)) { best = v; bi = i; }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_down(best, off, 64);
            const int oi = __shfl_down(bi, off, 64);
            if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
        }
        if ((tid & 63) == 0) { wv[tid >> 6] = best; wi[tid >> 6] = bi; }
        __syncthreads();
        if (tid == 0) {
            for (int w = 1; w < kGjThreads / 64; ++w)
                if (wv[w] > best || (wv[w] == best && wi[w] < bi)) { best = wv[w]; bi = wi[w]; }
            piv[k] = bi;
            if (!(best > 0.0)) bad = 1;
        }
        __syncthreads();
        if (bad) {
            if (tid == 0) atomicOr(fail, 2);
            return;
        }
        const int p = piv[k];
        if (p != k)
            for (int j = tid; j < n; j += kGjThreads) {
                const double tmp = M[(long)k * n + j];
                M[(long)k * n + j] = M[(long)p * n + j];
                M[(long)p * n + j] = tmp;
            }
        __syncthreads();
        const double r = 1.0 / M[(long)k * n + k];
        for (int i = tid; i < n; i += kGjThreads) fcol[i] = M[(long)i * n + k];
        __syncthreads();
        for (int j = tid; j < n; j += kGjThreads) M[(long)k * n + j] = (j == k ? 1.0 : M[(long)k * n + j]) * r;
        __syncthreads();
        for (long q = tid; q < nn; q += kGjThreads) {
            const int i = (int)(q / n), j = (int)(q - (long)i * n);
            if (i == k) continue;
            M[q] = (j == k ? 0.0 : M[q]) - M[(long)k * n + j] * fcol[i];
        }
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; --k) {
        const int p = piv[k];
        if (p != k)
            for (int i = tid; i < n; i += kGjThreads) {
                const double tmp = M[(long)i * n + k];
                M[(long)i * n + k] = M[(long)i * n + p];
                M[(long)i * n + p] = tmp;
            }
        __syncthreads();
    }
}

// x = M^-1 b on the coarsest level: one lane per row, columns in order
__global__ void k_amg_dense_apply(int n, const double* __restrict__ Minv, const double* __restrict__ b,
                                  double* __restrict__ x, const int* __restrict__ run)
{
    if (run && !*run) return;
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const double* row = Minv + (long)i * n;
    double s = row[0] * b[0];
    for (int j = 1; j < n; ++j) s += row[j] * b[j];
    x[i] = s;
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

AmgHierarchy::AmgHierarchy(int nx, int ny, int nz, const AmgParams& p, hipStream_t s) : p_(p)
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
        const bool last = l + 1 == dims.size();
        if (!last) {
            L.ax = p.agg[0]; L.ay = p.agg[1]; L.az = p.agg[2];
            L.dinv.alloc(9 * L.N);
            L.r.alloc(3 * L.N);
            L.tiles = std::make_unique<ColoredIlu>(L.nx, L.ny, L.nz, s);
        }
        if (l > 0) {
            L.A.alloc(63 * L.N);
            L.b.alloc(3 * L.N);
            L.x.alloc(3 * L.N);
        }
    }
    const long n = 3 * lv_.back()->N;
    dense_.alloc(n * n);
}

void AmgHierarchy::setup(const double* A0, int* fail, hipStream_t s)
{
    A0_ = A0;
    for (size_t l = 0; l < lv_.size(); ++l) {
        Level& L = *lv_[l];
        const double* A = matrix((int)l);
        if (l > 0) {
            const Level& F = *lv_[l - 1];
            const AmgGrid f = gridOf(F.nx, F.ny, F.nz), c = gridOf(L.nx, L.ny, L.nz);
            hipLaunchKernelGGL(k_amg_galerkin, dim3(blocksFor(9 * c.N, 256)), dim3(256), 0, s, f, c, F.ax, F.ay,
                               F.az, matrix((int)l - 1), L.A.p);
        }
        if (l + 1 < lv_.size()) launchDiaBlockInverse(L.N, A, L.dinv.p, fail, s);
    }
    const Level& C = *lv_.back();
    const AmgGrid c = gridOf(C.nx, C.ny, C.nz);
    const long n = 3 * C.N;
    hipLaunchKernelGGL(k_amg_dense_build, dim3(blocksFor(n * n, 256)), dim3(256), 0, s, c, matrix((int)lv_.size() - 1),
                       dense_.p);
    hipLaunchKernelGGL(k_amg_dense_invert, dim3(1), dim3(kGjThreads), 0, s, (int)n, dense_.p, fail);
    OM_HIP(hipGetLastError());
}

void AmgHierarchy::cycle(size_t l, const double* b, double* x, const int* run, hipStream_t s)
{
    Level& L = *lv_[l];
    if (l + 1 == lv_.size()) {
        const int n = (int)(3 * L.N);
        hipLaunchKernelGGL(k_amg_dense_apply, dim3(blocksFor(n, 256)), dim3(256), 0, s, n, dense_.p, b, x, run);
        return;
    }
    Level& C = *lv_[l + 1];
    const double* A = matrix((int)l);
    const AmgGrid f = gridOf(L.nx, L.ny, L.nz), c = gridOf(C.nx, C.ny, C.nz);
    for (int k = 0; k < p_.pre; ++k) L.tiles->gsSweep<double>(A, L.dinv.p, b, x, true, k == 0, s, run);
    hipLaunchKernelGGL(k_amg_residual, dim3(blocksFor(3 * f.N, 256)), dim3(256), 0, s, f, A, b, x, L.r.p, run);
    hipLaunchKernelGGL(k_amg_restrict, dim3(blocksFor(3 * c.N, 256)), dim3(256), 0, s, f, c, L.ax, L.ay, L.az,
                       L.r.p, C.b.p, run);
    cycle(l + 1, C.b.p, C.x.p, run, s);
    hipLaunchKernelGGL(k_amg_prolong, dim3(blocksFor(3 * f.N, 256)), dim3(256), 0, s, f, c, L.ax, L.ay, L.az, p_.damp,
                       C.x.p, x, run);
    for (int k = 0; k < p_.post; ++k) L.tiles->gsSweep<double>(A, L.dinv.p, b, x, false, false, s, run);
}

void AmgHierarchy::apply(const double* d, double* v, const int* run, hipStream_t s)
{
    if (!A0_) throw std::logic_error("AMG: apply before setup");
    cycle(0, d, v, run, s);
    OM_HIP(hipGetLastError());
}

}  // namespace omodels
