// ens_kernels.hip -- BiCGStab vector steps, SpMV and norms for an ensemble of
// reservoirs in single launches.  Blocks are cell chunks (kEnsChunk cells of
// one instance, table chunk -> (instance, chunk)); per-instance reductions
// sum the chunk partials of that instance in a fixed order, so an instance
// gets bit-identical scalars whether it is solved alone or in a batch.
// Instances whose solve has finished are masked by their `active` flag.
#include "ensemble.h"
#include "ens_ops.h"

namespace omodels {

namespace {

constexpr int kT = kEnsChunk;

template <class S> __device__ __forceinline__ S* vec(const EnsInst& I, int id);
template <> __device__ __forceinline__ double* vec<double>(const EnsInst& I, int id) { return I.dv[id]; }
template <> __device__ __forceinline__ float* vec<float>(const EnsInst& I, int id) { return I.fv[id]; }

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
    __shared__ double sm[kT / 64];
    v = IsMax ? waveMax(v) : waveSum(v);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) sm[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        r = sm[0];
        for (int q = 1; q < kT / 64; ++q) r = IsMax ? fmax(r, sm[q]) : r + sm[q];
    }
    __syncthreads();
    return r;
}
// sum / max of the partials of instance b (chunks [first[b], first[b+1]))
template <bool IsMax>
__device__ __forceinline__ double instReduce(const double* __restrict__ part, int c0, int c1)
{
    double v = 0.0;
    for (int k = c0 + (int)threadIdx.x; k < c1; k += kT) v = IsMax ? fmax(v, part[k]) : v + part[k];
    return blockReduce<IsMax>(v);
}

__device__ __forceinline__ void markFail(KrylovScalars* ks, int tag)
{
    if (!ks->fail_code) { ks->fail_code = 2; ks->fail_tag = tag; }
}

}  // namespace

// ---------------------------------------------------------------------------
template <class S>
__global__ __launch_bounds__(kT) void k_e_copy(const EnsInst* __restrict__ insts, const int2* __restrict__ table,
                                               const int* __restrict__ act, int dst, int src)
{
    const int2 t = table[blockIdx.x];
    if (!act[t.x]) return;
    const EnsInst& I = insts[t.x];
    const long c = (long)t.y * kT + threadIdx.x;
    if (c >= I.N) return;
    S* y = vec<S>(I, dst);
    if (src < 0) { y[3 * c] = S(0); y[3 * c + 1] = S(0); y[3 * c + 2] = S(0); }
    else {
        const S* x = vec<S>(I, src);
        y[3 * c] = x[3 * c]; y[3 * c + 1] = x[3 * c + 1]; y[3 * c + 2] = x[3 * c + 2];
    }
}

// Scalar products: the exact sum of the products (exact_sum.h), so a member's
// value does not depend on the batch or on the launch shape and equals the
// single-reservoir solver's.  A workgroup takes kEnsDotCells cells of one
// member (item = member, index among the member's workgroups, their count),
// adds its limbs to the member's device accumulators, and the member's last
// workgroup to finish rounds and evaluates the member's scalars.
// NACC = 1: a.b;  NACC = 2: a.a and a.b.  mode: 0 rho/beta, 1 denom/alpha,
// 2 omega (bicgstabsolver.hh:197-236, :287-292).
template <class S, int NACC>
__global__ __launch_bounds__(kT) void k_e_xdot(const EnsInst* __restrict__ insts, const int4* __restrict__ items,
    const int* __restrict__ act, int ia, int ib, ExactAcc* __restrict__ accAll,
    unsigned* __restrict__ tickets, KrylovScalars* __restrict__ ksAll, int mode, double eps, int tagOff)
{
    const int4 it = items[blockIdx.x];
    if (!act[it.x]) return;
    constexpr int nw = kT / 64;
    __shared__ unsigned long long sAcc[NACC * kAccLimbs];
    __shared__ unsigned sMeta[5];
    clearLimbs(sAcc, NACC, sMeta);
    __syncthreads();
    const EnsInst& I = insts[it.x];
    unsigned fl0 = 0u, fl1 = 0u;
    const S* a = vec<S>(I, ia);
    const S* b = vec<S>(I, ib);
    constexpr int kPass = kEnsDotCells / kT;     // cells per lane
    double v0[3 * kPass], v1[3 * kPass];
#pragma unroll
    for (int pass = 0; pass < kPass; ++pass) {
        const long c = (long)it.y * kEnsDotCells + (long)pass * kT + threadIdx.x;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            double x0 = 0.0, x1 = 0.0;
            if (c < I.N) {
                const double ai = (double)a[3 * c + e];
                if (NACC == 1) x0 = ai * (double)b[3 * c + e];
                else { x0 = ai * ai; x1 = ai * (double)b[3 * c + e]; }
            }
            v0[3 * pass + e] = x0;
            v1[3 * pass + e] = x1;
        }
    }
    accumulateBands<3 * kPass>(v0, sAcc, fl0);
    if (NACC == 2) accumulateBands<3 * kPass>(v1, sAcc + kAccLimbs, fl1);
    if (fl0 | fl1) atomicOr(&sMeta[0], fl0 | (fl1 << 3));
    __syncthreads();
    ExactAcc* acc = accAll + 2 * it.x;
    flushLimbs(sAcc, NACC, sMeta, acc);
    if (!lastToFinish(tickets + it.x, it.z, sMeta)) return;
    __shared__ long long sD[kAccLimbs];
    const double s0 = finishExact(&acc[0], sD, sMeta);
    const double s1 = NACC == 2 ? finishExact(&acc[1], sD, sMeta) : 0.0;
    if (threadIdx.x != 0) return;
    KrylovScalars* ks = ksAll + it.x;
    // failures carry the iteration's tag (tag_base = 2 it, advanced on the device)
    const int tag = ks->tag_base + tagOff;
    if (mode == 0) {
        ks->rho_i = s0;
        if (fabs(ks->rho) <= eps || fabs(ks->omega) <= eps) markFail(ks, tag);
        ks->beta = (s0 / ks->rho) * (ks->alpha / ks->omega);
        ks->rho = s0;
    } else if (mode == 1) {
        ks->denom = s0;
        if (fabs(s0) <= eps) markFail(ks, tag);
        ks->alpha = ks->rho / s0;
        if (fabs(ks->alpha) <= eps) markFail(ks, tag);
    } else {
        if (fabs(s0) <= eps) markFail(ks, tag);
        ks->omega = s1 / s0;
        if (fabs(ks->omega) <= eps) markFail(ks, tag);
    }
}

template <class S>
__global__ __launch_bounds__(kT) void k_e_update_p(const EnsInst* __restrict__ insts, const int2* __restrict__ table,
                                                   const int* __restrict__ act, const KrylovScalars* __restrict__ ks)
{
    const int2 t = table[blockIdx.x];
    if (!act[t.x]) return;
    const EnsInst& I = insts[t.x];
    const long c = (long)t.y * kT + threadIdx.x;
    if (c >= I.N) return;
    const S omega = (S)ks[t.x].omega, beta = (S)ks[t.x].beta;
    const S* r = vec<S>(I, kVr);
    const S* v = vec<S>(I, kVv);
    S* p = vec<S>(I, kVp);
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const long i = 3 * c + e;
        S tmp = v[i];
        tmp *= omega;
        tmp -= p[i];
        tmp *= -beta;
        S pn = r[i];
        pn += tmp;
        p[i] = pn;
    }
}

// x += alpha y; r -= alpha v; partials of max|r| and of (y != 0)
template <class S>
__global__ __launch_bounds__(kT) void k_e_update_xr(const EnsInst* __restrict__ insts, const int2* __restrict__ table,
    const int* __restrict__ act, const KrylovScalars* __restrict__ ks, double* __restrict__ part, int nChunks)
{
    const int2 t = table[blockIdx.x];
    if (!act[t.x]) return;
    const EnsInst& I = insts[t.x];
    const long c = (long)t.y * kT + threadIdx.x;
    double mx = 0.0, nz = 0.0;
    if (c < I.N) {
        const S alpha = (S)ks[t.x].alpha;
        const S* y = vec<S>(I, kVy);
        const S* v = vec<S>(I, kVv);
        S* x = vec<S>(I, kVx);
        S* r = vec<S>(I, kVr);
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const long i = 3 * c + e;
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
    }
    mx = blockReduce<true>(mx);
    nz = blockReduce<true>(nz);
    if (threadIdx.x == 0) { part[blockIdx.x] = mx; part[nChunks + blockIdx.x] = nz; }
}

__global__ __launch_bounds__(kT) void k_e_crit_final(const int* __restrict__ first, const int* __restrict__ act,
    const double* __restrict__ part, int nChunks, KrylovScalars* ksAll)
{
    const int b = blockIdx.x;
    if (!act[b]) return;
    const double mx = instReduce<true>(part, first[b], first[b + 1]);
    const double nz = instReduce<true>(part + nChunks, first[b], first[b + 1]);
    if (threadIdx.x != 0) return;
    ksAll[b].resnorm = mx;
    ksAll[b].stag_half = nz == 0.0 ? 1 : 0;
}

// x += omega z; r += -omega y; partial of (z != 0)
template <class S>
__global__ __launch_bounds__(kT) void k_e_update_xr2(const EnsInst* __restrict__ insts,
    const int2* __restrict__ table, const int* __restrict__ act, const KrylovScalars* __restrict__ ks,
    double* __restrict__ part)
{
    const int2 t = table[blockIdx.x];
    if (!act[t.x]) return;
    const EnsInst& I = insts[t.x];
    const long c = (long)t.y * kT + threadIdx.x;
    double nz = 0.0;
    if (c < I.N) {
        const S omega = (S)ks[t.x].omega;
        const S* z = vec<S>(I, kVz);
        const S* y = vec<S>(I, kVy);
        S* x = vec<S>(I, kVx);
        S* r = vec<S>(I, kVr);
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const long i = 3 * c + e;
            const S zi = z[i];
            x[i] += omega * zi;
            if (zi != S(0)) nz = 1.0;
            r[i] += -omega * y[i];
        }
    }
    nz = blockReduce<true>(nz);
    if (threadIdx.x == 0) part[blockIdx.x] = nz;
}

// the last kernel of an iteration: the stagnation verdict of the second half,
// then the next iteration's tag base (so a replayed graph needs no new
// arguments per iteration)
__global__ __launch_bounds__(kT) void k_e_stag_final(const int* __restrict__ first, const int* __restrict__ act,
    const double* __restrict__ part, KrylovScalars* ksAll, int tagOff)
{
    const int b = blockIdx.x;
    if (!act[b]) return;
    const double nz = instReduce<true>(part, first[b], first[b + 1]);
    if (threadIdx.x != 0) return;
    if (nz == 0.0 && ksAll[b].stag_full_tag < 0) ksAll[b].stag_full_tag = ksAll[b].tag_base + tagOff;
    ksAll[b].tag_base += 2;
}

// y = A x, A = FP64 Jacobian (useFloat = 0) or its FP32 copy
template <class S>
__global__ __launch_bounds__(kT) void k_e_spmv(const EnsInst* __restrict__ insts, const int2* __restrict__ table,
                                               const int* __restrict__ act, int ix, int iy)
{
    const int2 t = table[blockIdx.x];
    if (!act[t.x]) return;
    const EnsInst& I = insts[t.x];
    const long i = (long)t.y * kT + threadIdx.x;
    if (i >= I.N) return;
    const S* a;
    if constexpr (sizeof(S) == 8) a = (const S*)I.jac; else a = (const S*)I.Af;
    const S* x = vec<S>(I, ix);
    S* y = vec<S>(I, iy);
    S acc[3] = {S(0), S(0), S(0)};
    if (sizeof(S) == 4 && I.afDia) {
        // the structured layout: the present blocks in the same (column) order
        const int nx = I.G.nx, ny = I.G.ny, nz = I.G.nz;
        const long nxy = (long)nx * ny;
        const int ci = (int)(i % nx), cj = (int)((i / nx) % ny), ck = (int)(i / nxy);
        const long off[7] = {-nxy, -(long)nx, -1, 0, 1, (long)nx, nxy};
        const bool has[7] = {ck > 0, cj > 0, ci > 0, true, ci + 1 < nx, cj + 1 < ny, ck + 1 < nz};
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            if (!has[q]) continue;
            const long c = i + off[q];
            const S* blk = a + 9L * ((long)q * I.N + i);
            const S x0 = x[3 * c], x1 = x[3 * c + 1], x2 = x[3 * c + 2];
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                S s = blk[e * 3] * x0;
                s += blk[e * 3 + 1] * x1;
                s += blk[e * 3 + 2] * x2;
                acc[e] += s;
            }
        }
        y[3 * i] = acc[0]; y[3 * i + 1] = acc[1]; y[3 * i + 2] = acc[2];
        return;
    }
    const int k1 = I.rowPtr[i + 1];
    for (int k = I.rowPtr[i]; k < k1; ++k) {
        const long c = I.colIdx[k];
        const S* blk = a + 9L * k;
        const S x0 = x[3 * c], x1 = x[3 * c + 1], x2 = x[3 * c + 2];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            S s = blk[e * 3] * x0;
            s += blk[e * 3 + 1] * x1;
            s += blk[e * 3 + 2] * x2;
            acc[e] += s;
        }
    }
    y[3 * i] = acc[0]; y[3 * i + 1] = acc[1]; y[3 * i + 2] = acc[2];
}

// max |vec| per instance -> out[b]
template <class S>
__global__ __launch_bounds__(kT) void k_e_maxabs_partial(const EnsInst* __restrict__ insts,
    const int2* __restrict__ table, const int* __restrict__ act, int id, int useRes, double* __restrict__ part)
{
    const int2 t = table[blockIdx.x];
    if (!act[t.x]) return;
    const EnsInst& I = insts[t.x];
    const long c = (long)t.y * kT + threadIdx.x;
    double mx = 0.0;
    if (c < I.N) {
        const S* x = useRes ? (const S*)I.res : vec<S>(I, id);
#pragma unroll
        for (int e = 0; e < 3; ++e) mx = fmax(mx, fabs((double)x[3 * c + e]));
    }
    mx = blockReduce<true>(mx);
    if (threadIdx.x == 0) part[blockIdx.x] = mx;
}
__global__ __launch_bounds__(kT) void k_e_max_final(const int* __restrict__ first, const int* __restrict__ act,
                                                    const double* __restrict__ part, double* __restrict__ out)
{
    const int b = blockIdx.x;
    if (!act[b]) return;
    const double mx = instReduce<true>(part, first[b], first[b + 1]);
    if (threadIdx.x == 0) out[b] = mx;
}

// lu64 = J (row blocks of the chunk); with toFloat also Af = float(J)
// The block values of a chunk's rows are one contiguous range; it is copied
// with consecutive threads on consecutive doubles.
__device__ __forceinline__ void chunkValueRange(const EnsInst& I, int chunk, long& lo, long& hi)
{
    const long r0 = (long)chunk * kT;
    const long r1 = r0 + kT < I.N ? r0 + kT : I.N;
    lo = 9L * I.rowPtr[r0];
    hi = 9L * I.rowPtr[r1];
}
__global__ __launch_bounds__(kT) void k_e_copy_matrix(const EnsInst* __restrict__ insts,
    const int2* __restrict__ table, const int* __restrict__ act, int toFloat)
{
    const int2 t = table[blockIdx.x];
    if (!act[t.x]) return;
    const EnsInst& I = insts[t.x];
    long lo, hi;
    chunkValueRange(I, t.y, lo, hi);
    if (toFloat && I.afDia) {
        // the chunk's cells in the structured layout, slot by slot
        const long r0 = (long)t.y * kT;
        const int nc = (int)((r0 + kT < I.N ? r0 + kT : I.N) - r0);
        const int nx = I.G.nx, ny = I.G.ny, nz = I.G.nz;
        const long nxy = (long)nx * ny;
        for (int idx = threadIdx.x; idx < 7 * nc; idx += kT) {
            const int q = idx / nc;
            const long c = r0 + (idx - q * nc);
            const int ci = (int)(c % nx), cj = (int)((c / nx) % ny), ck = (int)(c / nxy);
            const bool has[7] = {ck > 0, cj > 0, ci > 0, true, ci + 1 < nx, cj + 1 < ny, ck + 1 < nz};
            float* out = I.Af + 9L * ((long)q * I.N + c);
            if (!has[q]) {
                for (int v = 0; v < 9; ++v) out[v] = 0.0f;
                continue;
            }
            long pos = I.rowPtr[c];
            for (int r = 0; r < q; ++r) pos += has[r] ? 1 : 0;
            for (int v = 0; v < 9; ++v) out[v] = (float)I.jac[9 * pos + v];
        }
    } else if (toFloat) {
        for (long k = lo + threadIdx.x; k < hi; k += kT) I.Af[k] = (float)I.jac[k];
    } else {
        for (long k = lo + threadIdx.x; k < hi; k += kT) I.lu64[k] = I.jac[k];
    }
}
// luf = float(lu64)
__global__ __launch_bounds__(kT) void k_e_factor_to_float(const EnsInst* __restrict__ insts,
    const int2* __restrict__ table, const int* __restrict__ act)
{
    const int2 t = table[blockIdx.x];
    if (!act[t.x]) return;
    const EnsInst& I = insts[t.x];
    long lo, hi;
    chunkValueRange(I, t.y, lo, hi);
    for (long k = lo + threadIdx.x; k < hi; k += kT) I.luf[k] = (float)I.lu64[k];
}

// mixed precision: fv[b] = float(r64); then x64 += double(fv[x]); r64 = b64 - J x64
__global__ __launch_bounds__(kT) void k_e_rhs_to_float(const EnsInst* __restrict__ insts,
    const int2* __restrict__ table, const int* __restrict__ act)
{
    const int2 t = table[blockIdx.x];
    if (!act[t.x]) return;
    const EnsInst& I = insts[t.x];
    const long c = (long)t.y * kT + threadIdx.x;
    if (c >= I.N) return;
    for (int e = 0; e < 3; ++e) I.fv[kVb][3 * c + e] = (float)I.dv[kVr64][3 * c + e];
}
__global__ __launch_bounds__(kT) void k_e_add_correction(const EnsInst* __restrict__ insts,
    const int2* __restrict__ table, const int* __restrict__ act)
{
    const int2 t = table[blockIdx.x];
    if (!act[t.x]) return;
    const EnsInst& I = insts[t.x];
    const long c = (long)t.y * kT + threadIdx.x;
    if (c >= I.N) return;
    for (int e = 0; e < 3; ++e) I.dv[kVx][3 * c + e] += (double)I.fv[kVx][3 * c + e];
}
__global__ __launch_bounds__(kT) void k_e_residual64(const EnsInst* __restrict__ insts,
    const int2* __restrict__ table, const int* __restrict__ act)
{
    const int2 t = table[blockIdx.x];
    if (!act[t.x]) return;
    const EnsInst& I = insts[t.x];
    const long i = (long)t.y * kT + threadIdx.x;
    if (i >= I.N) return;
    const double* x = I.dv[kVx];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int k = I.rowPtr[i]; k < I.rowPtr[i + 1]; ++k) {
        const long c = I.colIdx[k];
        const double* blk = I.jac + 9L * k;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            double s = blk[e * 3] * x[3 * c];
            s += blk[e * 3 + 1] * x[3 * c + 1];
            s += blk[e * 3 + 2] * x[3 * c + 2];
            acc[e] += s;
        }
    }
    for (int e = 0; e < 3; ++e) I.dv[kVr64][3 * i + e] = I.res[3 * i + e] - acc[e];
}

// ---------------------------------------------------------------------------
EnsOps::EnsOps(const EnsInst* insts, const int2* table, int nChunks, const int* first, int nInst, double* part,
               const EnsDots& dots, KrylovScalars* ks, hipStream_t s)
    : insts_(insts), table_(table), nChunks_(nChunks), first_(first), nInst_(nInst), part_(part), dots_(dots), ks_(ks),
      s_(s)
{}

#define EGRID dim3(nChunks_), dim3(kT), 0, s_
#define IGRID dim3(nInst_), dim3(kT), 0, s_
#define DGRID dim3(dots_.nItems), dim3(kT), 0, s_

template <class S> void EnsOps::copy(const int* act, int dst, int src)
{ if (nChunks_) hipLaunchKernelGGL((k_e_copy<S>), EGRID, insts_, table_, act, dst, src); }

template <class S> void EnsOps::dot(const int* act, int ia, int ib, int op, int tagOff)
{
    if (!nChunks_) return;
    const double eps = (double)std::numeric_limits<S>::min() * 1e10;
    hipLaunchKernelGGL((k_e_xdot<S, 1>), DGRID, insts_, dots_.items, act, ia, ib, dots_.acc, dots_.tickets, ks_, op,
                       eps, tagOff);
}

template <class S> void EnsOps::updateP(const int* act)
{ if (nChunks_) hipLaunchKernelGGL((k_e_update_p<S>), EGRID, insts_, table_, act, ks_); }

template <class S> void EnsOps::updateXR(const int* act)
{
    if (!nChunks_) return;
    hipLaunchKernelGGL((k_e_update_xr<S>), EGRID, insts_, table_, act, ks_, part_, nChunks_);
    hipLaunchKernelGGL(k_e_crit_final, IGRID, first_, act, part_, nChunks_, ks_);
}

template <class S> void EnsOps::dot2(const int* act, int tagOff)
{
    if (!nChunks_) return;
    const double eps = (double)std::numeric_limits<S>::min() * 1e10;
    hipLaunchKernelGGL((k_e_xdot<S, 2>), DGRID, insts_, dots_.items, act, kVy, kVr, dots_.acc, dots_.tickets, ks_, 2,
                       eps, tagOff);
}

template <class S> void EnsOps::updateXR2(const int* act, int tagOff)
{
    if (!nChunks_) return;
    hipLaunchKernelGGL((k_e_update_xr2<S>), EGRID, insts_, table_, act, ks_, part_);
    hipLaunchKernelGGL(k_e_stag_final, IGRID, first_, act, part_, ks_, tagOff);
}

template <class S> void EnsOps::spmv(const int* act, int ix, int iy)
{ if (nChunks_) hipLaunchKernelGGL((k_e_spmv<S>), EGRID, insts_, table_, act, ix, iy); }

template <class S> void EnsOps::maxAbs(const int* act, int id, int useRes, double* out)
{
    if (!nChunks_) return;
    hipLaunchKernelGGL((k_e_maxabs_partial<S>), EGRID, insts_, table_, act, id, useRes, part_);
    hipLaunchKernelGGL(k_e_max_final, IGRID, first_, act, part_, out);
}

void EnsOps::copyMatrix(const int* act, int toFloat)
{ if (nChunks_) hipLaunchKernelGGL(k_e_copy_matrix, EGRID, insts_, table_, act, toFloat); }
void EnsOps::factorToFloat(const int* act)
{ if (nChunks_) hipLaunchKernelGGL(k_e_factor_to_float, EGRID, insts_, table_, act); }
void EnsOps::rhsToFloat(const int* act)
{ if (nChunks_) hipLaunchKernelGGL(k_e_rhs_to_float, EGRID, insts_, table_, act); }
void EnsOps::addCorrection(const int* act)
{ if (nChunks_) hipLaunchKernelGGL(k_e_add_correction, EGRID, insts_, table_, act); }
void EnsOps::residual64(const int* act)
{ if (nChunks_) hipLaunchKernelGGL(k_e_residual64, EGRID, insts_, table_, act); }

#define EINST(S)                                                              \
    template void EnsOps::copy<S>(const int*, int, int);                     \
    template void EnsOps::dot<S>(const int*, int, int, int, int);            \
    template void EnsOps::updateP<S>(const int*);                            \
    template void EnsOps::updateXR<S>(const int*);                           \
    template void EnsOps::dot2<S>(const int*, int);                          \
    template void EnsOps::updateXR2<S>(const int*, int);                     \
    template void EnsOps::spmv<S>(const int*, int, int);                     \
    template void EnsOps::maxAbs<S>(const int*, int, int, double*);
EINST(double)
EINST(float)

}  // namespace omodels
