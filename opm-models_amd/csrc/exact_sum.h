// exact_sum.h -- dot products whose result does not depend on the summation
// order: the products are added exactly into a long fixed-point accumulator
// and the sum is rounded once, to the nearest double (ties to even).
//
// The reference sums sequentially (OverlappingScalarProduct::dot,
// overlappingscalarproduct.hh:58-70, a plain loop over the entries); that
// result depends on the order of the additions, so a parallel reduction can
// only reproduce it by fixing the order.  The exact sum is the value the
// sequential loop approximates, it is the same for every partition of the
// work (one reservoir or a member of an ensemble, any grid size), and the CPU
// restatement computes the same value.
//
// Representation.  A finite double is m * 2^(pos - 1074) with an integer
// m < 2^53 and 0 <= pos <= 2045.  An accumulator holds kAccLimbs signed
// 64-bit limbs, limb j weighing 2^(32 j - 1074).  A product is cut at the
// limb boundary below it (k = pos / 32): the integer m * 2^(pos mod 32) <
// 2^85 goes as three chunks below 2^32 into limbs k, k+1, k+2.  Carries are
// deferred: with fewer than 2^28 products no limb can overflow.
//
// Reduction.  A wave splits each product error-free into four bands fixed
// by the wave's largest product (q = (sigma + r) - sigma, r -= q, for
// sigma = 2^(e_max + 10), 2^(e_max - 33), ...): every band's parts are
// multiples of its unit and their sum over the wave stays below 2^53 units,
// so the bands add exactly in double; a product with bits below the last
// band (2^-119 and less of the largest) adds its remainder to LDS directly.
// The wave's four band sums go into the workgroup's limbs in LDS (waves whose
// largest product is outside [2^-700, 2^1000] use a register window of
// integer limbs instead).  Each workgroup then adds its nonzero limbs to the
// device accumulator with device-scope atomics; the last workgroup to finish
// (a ticket counter) reads the accumulator, rounds, and clears it and the
// ticket for the next use.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace omodels {

constexpr int kAccLimbs = 68;      // 32-bit digits from 2^-1074; the largest product ends in digit 66

// The device accumulator of one sum: limbs with deferred carries and the
// non-finite products seen (1 NaN, 2 +inf, 4 -inf); zero between products.
struct ExactAcc {
    unsigned long long limb[kAccLimbs];
    unsigned int flags, pad_;
};

// One product placed in the accumulator: |p| = (hi * 2^64 + lo) * 2^(32 k - 1074)
// with hi * 2^64 + lo < 2^85; k < 0 for a zero or non-finite product.
struct ProductPiece {
    unsigned long long lo, hi;
    int k, neg;
};

__host__ __device__ __forceinline__ ProductPiece makePiece(double p, unsigned& flags)
{
    ProductPiece q;
    union { double d; unsigned long long u; } cv;
    cv.d = p;
    const unsigned long long b = cv.u;
    const int be = (int)((b >> 52) & 0x7ffu);
    unsigned long long man = b & 0xFFFFFFFFFFFFFull;
    q.neg = (int)(b >> 63);
    q.lo = 0ull; q.hi = 0ull; q.k = -1;
    if (be == 0x7ff) { flags |= man ? 1u : (q.neg ? 4u : 2u); return q; }
    if (be == 0 && man == 0ull) return q;
    int pos = 0;
    if (be) { man |= 1ull << 52; pos = be - 1; }
    q.k = pos >> 5;
    const int sh = pos & 31;
    q.lo = man << sh;
    q.hi = sh ? (man >> (64 - sh)) : 0ull;
    return q;
}

// The correctly rounded value of the accumulator.  d: the limbs as signed
// integers (overwritten), nonzero at most in [lo, hi]; flags as in ExactAcc.
__host__ __device__ inline double exactRound(long long* d, int lo, int hi, unsigned flags)
{
    union { double d; unsigned long long u; } cv;
    if ((flags & 1u) || (flags & 6u) == 6u) { cv.u = 0x7FF8000000000000ull; return cv.d; }
    if (flags & 2u) { cv.u = 0x7FF0000000000000ull; return cv.d; }
    if (flags & 4u) { cv.u = 0xFFF0000000000000ull; return cv.d; }
    if (lo > hi) return 0.0;
    // carries: digits in [0, 2^32), the sign in the carry out of digit hi + 1
    const int top = hi + 1 < kAccLimbs ? hi + 1 : kAccLimbs - 1;
    if (top > hi) d[top] = 0;
    long long carry = 0;
    for (int j = lo; j <= top; ++j) {
        const long long t = d[j] + carry;
        d[j] = t & 0xffffffffll;
        carry = t >> 32;
    }
    const bool neg = carry < 0;
    if (neg) {   // magnitude: two's complement of the digits
        long long c = 1;
        for (int j = lo; j <= top; ++j) {
            const long long t = (~d[j] & 0xffffffffll) + c;
            d[j] = t & 0xffffffffll;
            c = t >> 32;
        }
    }
    int h = top;
    while (h >= lo && d[h] == 0) --h;
    if (h < lo) return 0.0;
    const unsigned long long sgn = neg ? (1ull << 63) : 0ull;
    int lz = 0;
    for (unsigned long long v = (unsigned long long)d[h]; !(v & 0x80000000ull); v <<= 1) ++lz;
    int b = 32 * h + 31 - lz;       // leading bit, in units of 2^-1074
    if (b < 53) {                   // subnormal or smallest normal: exact
        unsigned long long v = 0ull;
        for (int j = h; j >= lo; --j) v = (v << 32) | (unsigned long long)d[j];
        v <<= 32 * lo;
        cv.u = sgn | v;
        return cv.d;
    }
    // the 64 bits ending at the leading bit, and whether anything lies below them
    auto dig = [&](int j) -> unsigned long long { return (j >= lo && j <= top) ? (unsigned long long)d[j] : 0ull; };
    const int s = 32 - lz;          // bits of digit h - 2 below the window
    const unsigned long long hiw = dig(h), midw = dig(h - 1), low = dig(h - 2);
    // window = (hiw * 2^64 + midw * 2^32 + low) >> s, s in [1, 32]
    const unsigned long long w = (hiw << (64 - s)) | (midw << (32 - s)) | (low >> s);
    bool sticky = (low & ((1ull << s) - 1ull)) != 0ull;
    for (int j = h - 3; j >= lo && !sticky; --j) sticky = d[j] != 0;
    unsigned long long mant = w >> 11;
    const bool rbit = (w >> 10) & 1ull;
    const bool rest = (w & 0x3ffull) != 0ull || sticky;
    if (rbit && (rest || (mant & 1ull))) ++mant;
    if (mant >> 53) { mant >>= 1; ++b; }
    const int biased = b - 51;
    if (biased >= 2047) { cv.u = sgn | 0x7FF0000000000000ull; return cv.d; }
    cv.u = sgn | ((unsigned long long)biased << 52) | (mant & 0xFFFFFFFFFFFFFull);
    return cv.d;
}

#ifdef __HIPCC__
constexpr int kWindow = 8;      // limbs of a wave's register window (256 bits)

// Adds U products per lane of a wave to the workgroup's limbs sAcc (LDS,
// kAccLimbs).  Every lane of the wave must call it (absent products: 0.0).
template <int U>
__device__ __forceinline__ void accumulateWave(const double (&p)[U], unsigned long long* sAcc, unsigned& flags)
{
    int k[U];
    unsigned long long c0[U], c1[U], c2[U];
    int kmax = -1;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const ProductPiece q = makePiece(p[u], flags);
        k[u] = q.k;
        c0[u] = q.lo & 0xffffffffull; c1[u] = q.lo >> 32; c2[u] = q.hi;
        if (q.neg) { c0[u] = 0ull - c0[u]; c1[u] = 0ull - c1[u]; c2[u] = 0ull - c2[u]; }
        kmax = max(kmax, q.k);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) kmax = max(kmax, __shfl_xor(kmax, o, 64));
    if (kmax < 0) return;                        // wave-uniform: no nonzero finite product
    const int kb = kmax + 3 - kWindow;           // absolute limb of window limb 0
    unsigned long long acc[kWindow];
#pragma unroll
    for (int w = 0; w < kWindow; ++w) acc[w] = 0ull;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (k[u] < 0) continue;
        const int t = k[u] - kb;
        if (t >= 0) {
#pragma unroll
            for (int w = 0; w < kWindow; ++w)
                acc[w] += w == t ? c0[u] : (w == t + 1 ? c1[u] : (w == t + 2 ? c2[u] : 0ull));
        } else {                                 // far below the wave's largest product
            atomicAdd(&sAcc[k[u]], c0[u]);
            atomicAdd(&sAcc[k[u] + 1], c1[u]);
            atomicAdd(&sAcc[k[u] + 2], c2[u]);
        }
    }
#pragma unroll
    for (int w = 0; w < kWindow; ++w)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc[w] += __shfl_xor(acc[w], o, 64);
    const int lane = (int)(threadIdx.x & 63u);
    unsigned long long v = 0ull;
#pragma unroll
    for (int w = 0; w < kWindow; ++w) v = lane == w ? acc[w] : v;
    if (lane < kWindow && v && kb + lane >= 0) atomicAdd(&sAcc[kb + lane], v);
}

// one value's chunks straight into LDS limbs
__device__ __forceinline__ void addToLimbs(double v, unsigned long long* sAcc, unsigned& flags)
{
    const ProductPiece q = makePiece(v, flags);
    if (q.k < 0) return;
    unsigned long long c0 = q.lo & 0xffffffffull, c1 = q.lo >> 32, c2 = q.hi;
    if (q.neg) { c0 = 0ull - c0; c1 = 0ull - c1; c2 = 0ull - c2; }
    atomicAdd(&sAcc[q.k], c0);
    atomicAdd(&sAcc[q.k + 1], c1);
    if (c2) atomicAdd(&sAcc[q.k + 2], c2);
}

// Adds U products per lane of a wave to the workgroup's limbs sAcc (LDS) by
// the error-free band split.  Every lane of the wave must call it (absent
// products: 0.0); U <= 64.
template <int U>
__device__ __forceinline__ void accumulateBands(const double (&p)[U], unsigned long long* sAcc, unsigned& flags)
{
    int emax = -2000;
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (p[u] != 0.0 && isfinite(p[u])) emax = max(emax, ilogb(p[u]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) emax = max(emax, __shfl_xor(emax, o, 64));
    if (emax < -700 || emax > 1000) {            // no finite nonzero product, or extreme ones
        accumulateWave<U>(p, sAcc, flags);
        return;
    }
    constexpr int kBands = 4;
    double S[kBands];
#pragma unroll
    for (int i = 0; i < kBands; ++i) S[i] = 0.0;
    const double sigma0 = ldexp(1.0, emax + 10);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (!isfinite(p[u])) { addToLimbs(p[u], sAcc, flags); continue; }
        double r = p[u], sigma = sigma0;
#pragma unroll
        for (int i = 0; i < kBands; ++i) {
            const double q = (sigma + r) - sigma;
            S[i] += q;
            r -= q;
            sigma *= 0x1p-43;
        }
        if (r != 0.0) addToLimbs(r, sAcc, flags);
    }
#pragma unroll
    for (int i = 0; i < kBands; ++i)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) S[i] += __shfl_xor(S[i], o, 64);
    const int lane = (int)(threadIdx.x & 63u);
    if (lane < kBands) {
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < kBands; ++i) v = lane == i ? S[i] : v;
        addToLimbs(v, sAcc, flags);
    }
}

// The workgroup's limbs of nAcc sums (LDS, nAcc * kAccLimbs) and their flags
// (3 bits each) in sMeta[0]; sMeta[1..4] scratch.
__device__ __forceinline__ void clearLimbs(unsigned long long* sAcc, int nAcc, unsigned* sMeta)
{
    for (int t = (int)threadIdx.x; t < nAcc * kAccLimbs; t += (int)blockDim.x) sAcc[t] = 0ull;
    if (threadIdx.x == 0) sMeta[0] = 0u;
}

// Adds the workgroup's limbs of nAcc sums to the device accumulators g[0..nAcc)
// (after the barrier that ends the accumulation).
__device__ __forceinline__ void flushLimbs(const unsigned long long* sAcc, int nAcc, const unsigned* sMeta,
                                           ExactAcc* g)
{
    for (int t = (int)threadIdx.x; t < nAcc * kAccLimbs; t += (int)blockDim.x) {
        const unsigned long long v = sAcc[t];
        if (v) __hip_atomic_fetch_add(&g[t / kAccLimbs].limb[t % kAccLimbs], v, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
    }
    if (threadIdx.x == 0 && sMeta[0])
        for (int a = 0; a < nAcc; ++a) {
            const unsigned f = (sMeta[0] >> (3 * a)) & 7u;
            if (f) __hip_atomic_fetch_or(&g[a].flags, f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
}

// After every thread's atomics: true in every thread of the workgroup that
// finishes last among nBlocks; that workgroup resets the ticket.
__device__ __forceinline__ bool lastToFinish(unsigned* ticket, int nBlocks, unsigned* sMeta)
{
    __builtin_amdgcn_s_waitcnt(0x0F70);     // this wave's atomics acknowledged
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const bool last = t == (unsigned)nBlocks - 1u;
        if (last) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sMeta[1] = last ? 1u : 0u;
    }
    __syncthreads();
    return sMeta[1] != 0u;
}

// In the last workgroup: the rounded value of the device accumulator g, which
// is cleared for the next use.  sD: LDS of kAccLimbs long longs.
__device__ __forceinline__ double finishExact(ExactAcc* g, long long* sD, unsigned* sMeta)
{
    if (threadIdx.x < kAccLimbs) {
        const unsigned long long v =
            __hip_atomic_load(&g->limb[threadIdx.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sD[threadIdx.x] = (long long)v;
        if (v) __hip_atomic_store(&g->limb[threadIdx.x], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (threadIdx.x == 0) {
        sMeta[2] = __hip_atomic_load(&g->flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (sMeta[2]) __hip_atomic_store(&g->flags, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    __shared__ double sR;
    if (threadIdx.x == 0) {
        int lo = kAccLimbs, hi = -1;
        for (int t = 0; t < kAccLimbs; ++t)
            if (sD[t]) { if (lo == kAccLimbs) lo = t; hi = t; }
        sR = exactRound(sD, lo, hi, sMeta[2]);
    }
    __syncthreads();
    return sR;
}
#endif

}  // namespace omodels
