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
// m < 2^53 and 0 <= pos <= 2045.  The accumulator holds kAccLimbs signed
// 64-bit limbs, limb j weighing 2^(32 j - 1074); a product is cut at the
// limb boundary below it, so it adds a value below 2^32 to each of at most
// four limbs.  Carries are deferred: with fewer than 2^28 products per
// accumulator no limb can overflow.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace omodels {

constexpr int kAccLimbs = 68;      // 32-bit digits from 2^-1074; the largest product ends in digit 66

struct ExactAcc {
    unsigned long long limb[kAccLimbs];   // two's-complement limbs, carries deferred
    unsigned int flags;                    // non-finite products seen: 1 NaN, 2 +inf, 4 -inf
    unsigned int pad_;
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
// Adds the pieces of a wave (Q per lane) into the workgroup's limbs in LDS.
// Lanes whose pieces share a limb index are summed exactly in 128 bits across
// the wave, then one lane adds the sum to four limbs.  Every lane of the wave
// must call this (absent entries are pieces with k < 0).
template <int Q>
__device__ __forceinline__ void accumulatePieces(const ProductPiece (&pc)[Q], unsigned long long* sAcc)
{
    unsigned pend = 0u;
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (pc[q].k >= 0) pend |= 1u << q;
    const int lane = (int)(threadIdx.x & 63u);
    for (;;) {
        const unsigned long long any = __ballot(pend != 0u);
        if (any == 0ull) break;
        const int leader = __ffsll((unsigned long long)any) - 1;
        int myk = 0;
#pragma unroll
        for (int q = Q - 1; q >= 0; --q)
            if ((pend >> q) & 1u) myk = pc[q].k;
        const int kk = __shfl(myk, leader, 64);
        unsigned long long lo = 0ull;
        long long hi = 0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            if (((pend >> q) & 1u) && pc[q].k == kk) {
                pend &= ~(1u << q);
                if (!pc[q].neg) {
                    const unsigned long long n = lo + pc[q].lo;
                    hi += (long long)pc[q].hi + (long long)(n < lo);
                    lo = n;
                } else {
                    const unsigned long long n = lo - pc[q].lo;
                    hi -= (long long)pc[q].hi + (long long)(lo < pc[q].lo);
                    lo = n;
                }
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long olo = __shfl_xor(lo, o, 64);
            const long long ohi = __shfl_xor(hi, o, 64);
            const unsigned long long n = lo + olo;
            hi += ohi + (long long)(n < lo);
            lo = n;
        }
        if (lane == 0) {
            atomicAdd(&sAcc[kk], lo & 0xffffffffull);
            atomicAdd(&sAcc[kk + 1], lo >> 32);
            atomicAdd(&sAcc[kk + 2], (unsigned long long)hi & 0xffffffffull);
            atomicAdd(&sAcc[kk + 3], (unsigned long long)(hi >> 32));
        }
    }
}

// Workgroup-level accumulation: clear the LDS limbs, ... accumulate ..., flush
// the nonzero limbs into the device accumulator with device-scope atomics.
__device__ __forceinline__ void clearLimbs(unsigned long long* sAcc, int nAcc, unsigned* sFlags)
{
    for (int t = (int)threadIdx.x; t < nAcc * kAccLimbs; t += (int)blockDim.x) sAcc[t] = 0ull;
    if (threadIdx.x == 0) *sFlags = 0u;
}
__device__ __forceinline__ void flushLimbs(const unsigned long long* sAcc, int nAcc, const unsigned* sFlags,
                                           ExactAcc* g)
{
    for (int t = (int)threadIdx.x; t < nAcc * kAccLimbs; t += (int)blockDim.x) {
        const unsigned long long v = sAcc[t];
        if (v) __hip_atomic_fetch_add(&g[t / kAccLimbs].limb[t % kAccLimbs], v, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
    }
    if (threadIdx.x == 0 && *sFlags)
        for (int a = 0; a < nAcc; ++a)
            __hip_atomic_fetch_or(&g[a].flags, (*sFlags >> (3 * a)) & 7u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Reads and clears the device accumulator g and returns its rounded value in
// every thread.  The block is exactly one wave of 64 threads; sD: LDS scratch
// of kAccLimbs long longs.
__device__ __forceinline__ double finishExact(ExactAcc* g, long long* sD)
{
    const int lane = (int)threadIdx.x;
    const unsigned long long a0 = __hip_atomic_load(&g->limb[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long a1 = 0ull;
    if (lane + 64 < kAccLimbs)
        a1 = __hip_atomic_load(&g->limb[lane + 64], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned flags = __hip_atomic_load(&g->flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (a0) __hip_atomic_store(&g->limb[lane], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (a1) __hip_atomic_store(&g->limb[lane + 64], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (lane == 0 && flags) __hip_atomic_store(&g->flags, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long nz0 = __ballot(a0 != 0ull), nz1 = __ballot(a1 != 0ull);
    sD[lane] = (long long)a0;
    if (lane + 64 < kAccLimbs) sD[lane + 64] = (long long)a1;
    int lo = kAccLimbs, hi = -1;
    if (nz0) { lo = __ffsll(nz0) - 1; hi = 63 - __clzll((long long)nz0); }
    if (nz1) {
        if (lo == kAccLimbs) lo = 64 + __ffsll(nz1) - 1;
        hi = 64 + 63 - __clzll((long long)nz1);
    }
    __syncthreads();
    __shared__ double sR;
    if (lane == 0) sR = exactRound(sD, lo, hi, flags);
    __syncthreads();
    return sR;
}
#endif

}  // namespace omodels
