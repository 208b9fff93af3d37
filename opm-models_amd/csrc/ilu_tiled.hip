// ilu_tiled.hip -- block ILU(0) factorization and triangular sweeps of the
// 7-point structured-grid pattern as persistent, tile-wavefront kernels, for
// one reservoir or for an ensemble of reservoirs in the same launch.
//
// Semantics are those of the level-scheduled kernels (dune-istl SeqILU<.,.,.,0>,
// natural ordering, istlpreconditionerwrappers.hh:151-190): same rows, same
// per-row operation order.  Only the schedule differs.
//
// Schedule: each grid is cut into BX x BY x BZ tiles (at most 240 cells).  A
// tile depends on its -x, -y and -z neighbour tiles (+x, +y, +z for the
// backward sweep) of the same instance; tiles of all instances are handed out
// by one atomic ticket in order of tile level (TI + TJ + TK), so every
// dependency of a tile holds a smaller ticket.  Only a running workgroup draws
// tickets, and it draws its next ticket while working on a smaller one, so the
// holder of the lowest unfinished ticket is always running and its
// dependencies are finished: it proceeds, and by induction every wait ends,
// whether or not all workgroups of the launch are resident at once.  The
// kernels are therefore launched as ordinary kernels (a cooperative launch
// would guarantee co-residency the schedule does not need, at the price of an
// idle gap of about 20 us around every launch).  Waits are also bounded and
// report an error instead of spinning forever.
//
// The factorization (once per Newton iteration) uses four lanes per cell with
// the tile's BX+BY+BZ-2 local levels separated by workgroup barriers, and a
// tile publishes a done flag after its inverted pivots are written through to
// the device-coherent level.  The sweeps, applied twice per solver iteration,
// come in two forms: batched like the factorization, or streaming: each face value
// goes to its neighbour tile as soon as it is computed, and inside a tile the
// levels follow each other through per-level completion counts in LDS rather
// than workgroup barriers, so a tile runs a few cell levels behind its
// upstream tiles instead of a whole tile behind them.
#include <algorithm>
#include <cmath>
#include <array>
#include <map>
#include <mutex>
#include <stdexcept>
#include <cstring>

#include "ensemble.h"
#include "ilu_tiled.h"

namespace omodels {

namespace {

constexpr int kPadMax = 704;            // bound of (BX+1)(BY+1)(BZ+1) over all tile shapes
constexpr long kSpinLimit = 1L << 22;   // a few seconds of polling

// Cross-tile exchange.  Every value that another tile reads while this kernel
// runs (sweep outputs, inverted pivots) and every done-flag is written and read
// with device-scope relaxed atomics: on gfx950 these are write-through stores
// and L2-bypassing loads (the sc1 forms), coherent across the XCDs without any
// cache-wide maintenance.  A tile publishes its flag only after each of its
// threads has seen its own write-through stores acknowledged (vmcnt(0)) and
// the workgroup has met at a barrier; a dependent tile issues its exchange
// loads only after its poll has returned the flag.  The agent-scope
// release/acquire pair that would express the same order in the language
// model costs a write-back and an invalidation of the whole L2 of the XCD per
// tile (buffer_wbl2 / buffer_inv), which dominated the sweep time; it is not
// needed because no exchanged value is ever accessed non-atomically while the
// kernel runs.
//
// Every global-memory access of these kernels goes through global-address-space
// pointers: generic (flat) accesses would also count against the LDS wait
// counter, so each LDS wait would wait for the wave's global stores as well.
template <class S> using GS = __attribute__((address_space(1))) S;
template <class S> __device__ __forceinline__ GS<S>* toGlobal(S* p) { return (GS<S>*)p; }
template <class S> __device__ __forceinline__ const GS<S>* toGlobal(const S* p) { return (const GS<S>*)p; }

__device__ __forceinline__ int loadRelaxed(const int* p)
{
    return __hip_atomic_load(const_cast<GS<int>*>(toGlobal(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class S>
__device__ __forceinline__ S ldCoh(const S* p)
{
    return __hip_atomic_load(const_cast<GS<S>*>(toGlobal(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class S>
__device__ __forceinline__ void stCoh(S* p, S v)
{
    __hip_atomic_store(toGlobal(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Sweep exchange.  A value that a neighbour tile needs is written to its own
// slot of a per-direction exchange array, which holds a signalling-NaN
// sentinel while empty; the consumer polls the slot (one aligned 8- or
// 4-byte load, atomic) until it differs from the sentinel, takes the value
// and writes the sentinel back for the next sweep.  Arithmetic never yields a
// signalling NaN; a value that is one (a copied input) is sent as a quiet NaN.
constexpr unsigned long long kEmpty64 = 0x7FF4A5A5A5A5A5A5ULL;
constexpr unsigned int kEmpty32 = 0x7FA5A5A5U;
__device__ __forceinline__ bool isEmpty(double v) { return (unsigned long long)__double_as_longlong(v) == kEmpty64; }
__device__ __forceinline__ bool isEmpty(float v) { return (unsigned int)__float_as_int(v) == kEmpty32; }
template <class S> __device__ __forceinline__ S emptySlot();
template <> __device__ __forceinline__ double emptySlot<double>() { return __longlong_as_double((long long)kEmpty64); }
template <> __device__ __forceinline__ float emptySlot<float>() { return __int_as_float((int)kEmpty32); }
template <class S>
__device__ __forceinline__ void sendSlot(S* slot, S v)
{
    stCoh(slot, isEmpty(v) ? S(NAN) : v);
}
// wait for a neighbour's value; bounded, and cut short once the launch has an error
template <class S>
__device__ __forceinline__ S receiveSlot(S* slot, int* err)
{
    long spins = 0;
    S v = ldCoh(slot);
    while (isEmpty(v)) {
        if (++spins > kSpinLimit || ((spins & 255) == 0 && loadRelaxed(err))) { atomicOr(err, 1); break; }
        __builtin_amdgcn_s_sleep(1);
        v = ldCoh(slot);
    }
    stCoh(slot, emptySlot<S>());
    return v;
}

// wait until every vector memory operation of this wave has completed
// (s_waitcnt vmcnt(0), expcnt and lgkmcnt left at their maxima)
__device__ __forceinline__ void waitStores() { __builtin_amdgcn_s_waitcnt(0x0F70); }

__device__ __forceinline__ void tileOf(const TileGeom& g, int tile, int& TI, int& TJ, int& TK)
{
    TI = tile % g.TX;
    TJ = (tile / g.TX) % g.TY;
    TK = tile / (g.TX * g.TY);
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
        // once any wait of this launch has run out (err set), the remaining
        // tiles stop waiting too, so a broken protocol ends within one limit
        for (int q = 0; q < nd && !loadRelaxed(err); ++q) {
            long spins = 0;
            while (loadRelaxed(done + deps[q]) != epoch) {
                if (++spins > kSpinLimit || ((spins & 1023) == 0 && loadRelaxed(err))) { atomicOr(err, 1); break; }
                __builtin_amdgcn_s_sleep(2);
            }
        }
    }
    __syncthreads();
}

// Tickets: thread 0 draws the next ticket while the current tile is worked
// on and hands it to the workgroup at the end of the tile.  Holding one drawn
// ticket while finishing a smaller one keeps the progress argument intact:
// the holder of the smallest unfinished ticket is always working on it.
__device__ __forceinline__ int drawTicket(unsigned long long* counter, long long base)
{
    return (int)((long long)atomicAdd(counter, 1ULL) - base);
}
__device__ __forceinline__ int shareTicket(int t, int* sTicket)
{
    if (threadIdx.x == 0) *sTicket = t;
    __syncthreads();
    return *sTicket;
}

// The last workgroup to leave a launch resets the ticket counter, so every
// launch (eager or replayed from a captured graph) draws tickets from zero:
// all workgroups have drawn their final ticket before they leave.
__device__ __forceinline__ void leaveLaunch(unsigned long long* counter, unsigned int* exits)
{
    if (threadIdx.x == 0 && atomicAdd(exits, 1u) == gridDim.x - 1) {
        atomicExch(counter, 0ULL);
        atomicExch(exits, 0u);
    }
}

__device__ __forceinline__ void publish(int* done, int tile, int epoch)
{
    waitStores();
    __syncthreads();
    if (threadIdx.x == 0) stCoh(done + tile, epoch);
}

__device__ __forceinline__ bool isActive(const IluInst& I) { return I.active == nullptr || *I.active != 0; }

// A launch for one instance that is inactive returns at once: no workgroup
// draws a ticket, the last one to leave resets the counter as usual.
__device__ __forceinline__ bool skipLaunch(const IluInst* insts, int nInst, unsigned long long* counter,
                                           unsigned int* exits)
{
    if (nInst != 1 || isActive(insts[0])) return false;
    leaveLaunch(counter, exits);
    return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// Sweeps.  A sweep workgroup has 16 waves.  Waves 0..14 compute, four lanes
// per tile cell (at most kStreamCells cells) with the cells taken in level
// order: lanes 4r..4r+3 handle the r-th cell of the tile's level sequence,
// one lane per block row (the fourth lane repeats the third row and stores
// nothing), and keep their rows of the cell's blocks and the right-hand side
// in registers.  Wave 15 receives the neighbour tiles' face values.
//
// Inside a tile the sweep is a dataflow, as it is across tiles: every value
// slot of the tile's LDS image (its cells and the ghost layer) starts as the
// exchange sentinel; a lane computes its cell once the nine values of its
// three dependency cells have left the sentinel, then writes its row's result
// to LDS and, for a cell on a face, straight into the exchange slot of the
// dependent tile.  The receiving wave polls all exchange slots of the tile's
// ghost layer together and moves every value that has arrived into LDS.  A
// tile thus starts as soon as its first ghost values arrive and follows its
// upstream tiles cell level by cell level, with no workgroup barrier between
// levels.  Per row the operations and their order are those of dune's block
// mmv.
constexpr int kSweepThreads = 1024;
constexpr int kStreamCells = 240;                // compute cells per tile: 15 waves of 16 cells
constexpr int kRecvWave = kStreamCells / 16;     // the receiving wave
constexpr int kRecvPerLane = 8;                  // ghost values per receiving lane
constexpr int kRecvDepth = 1;                    // loads in flight per ghost value (two measured slower: the extra polls load the fabric)
constexpr int kMaxTileLevels = 256;              // cell levels of a tile, BX + BY + BZ - 2 < 256

__device__ __forceinline__ int tri3(int m) { return m < 0 ? 0 : (m + 1) * (m + 2) * (m + 3) / 6; }
// number of cells (a, b, c) of a BX x BY x BZ box with a + b + c < l
__device__ __forceinline__ int cellsBelow(int l, int BX, int BY, int BZ)
{
    const int m = l - 1;
    return tri3(m) - tri3(m - BX) - tri3(m - BY) - tri3(m - BZ) + tri3(m - BX - BY) + tri3(m - BX - BZ) +
           tri3(m - BY - BZ) - tri3(m - BX - BY - BZ);
}
// position of (a, b, c) among the cells of its level a + b + c, ordered by (c, b)
__device__ __forceinline__ int rankInLevel(int a, int b, int c, int BX, int BY)
{
    const int l = a + b + c;
    int r = 0;
    for (int cp = 0; cp < c; ++cp) {
        const int lo = max(0, l - cp - BX + 1), hi = min(BY - 1, l - cp);
        r += max(0, hi - lo + 1);
    }
    return r + b - max(0, l - c - BX + 1);
}

// one word per tile position: local coordinates, which dependency neighbours
// exist, and whether the cell lies inside the grid
__device__ __forceinline__ int packCell(int li, int lj, int lk, bool n0, bool n1, bool n2, bool valid)
{
    return li | (lj << 8) | (lk << 16) | ((int)n0 << 27) | ((int)n1 << 28) | ((int)n2 << 29) | ((int)valid << 30);
}

// value of lane Q of this lane's quad (all four lanes of the quad are active)
template <int Q>
__device__ __forceinline__ double quadLane(double v)
{
    constexpr int ctrl = Q | (Q << 2) | (Q << 4) | (Q << 6);
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
template <int Q>
__device__ __forceinline__ float quadLane(float v)
{
    constexpr int ctrl = Q | (Q << 2) | (Q << 4) | (Q << 6);
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xF, 0xF, false));
}

template <class S>
__device__ __forceinline__ S rowDot(const S* b, const S* x)
{
    S s = b[0] * x[0];
    s += b[1] * x[1];
    s += b[2] * x[2];
    return s;
}

// LDS values polled by other waves: single aligned accesses the compiler keeps
template <class S>
__device__ __forceinline__ S ldsLoad(const S* p)
{
    return __hip_atomic_load(const_cast<S*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class S>
__device__ __forceinline__ void ldsStore(S* p, S v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One tile's rows of y = A v inside the backward sweep, once the sweep has
// stored the results of the tile and of its six neighbours (their flags carry
// this launch's epoch; a tile may finish before its +x, +y, +z neighbours have
// stored all their results, since it only waits for their faces).  The v values of the tile and its
// one-cell box around it are staged in LDS with device-scope loads (other
// workgroups stored them), then lane u takes row u % 3 of the tile's cell
// u / 3 with the blocks of the row in column order; per row the operation
// order is that of k_spmv3, so the product is bit-identical to the separate
// kernel's.
template <class S>
__device__ __forceinline__ void spmvTile(const IluInst& I, const SpmvFuse& fz, int tile, int ep, int* err, S* box)
{
    const TileGeom& g = I.g;
    int TI, TJ, TK;
    tileOf(g, tile, TI, TJ, TK);
    if (threadIdx.x == 0) {
        int deps[7];
        int nd = 0;
        deps[nd++] = tile;
        if (TI > 0) deps[nd++] = tile - 1;
        if (TJ > 0) deps[nd++] = tile - g.TX;
        if (TK > 0) deps[nd++] = tile - g.TX * g.TY;
        if (TI + 1 < g.TX) deps[nd++] = tile + 1;
        if (TJ + 1 < g.TY) deps[nd++] = tile + g.TX;
        if (TK + 1 < g.TZ) deps[nd++] = tile + g.TX * g.TY;
        for (int q = 0; q < nd && !loadRelaxed(err); ++q) {
            long spins = 0;
            while (loadRelaxed(fz.tileDone + deps[q]) != ep) {
                if (++spins > kSpinLimit || ((spins & 1023) == 0 && loadRelaxed(err))) { atomicOr(err, 1); break; }
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
    __syncthreads();
    const int BX2 = g.BX + 2, BXY2 = (g.BX + 2) * (g.BY + 2);
    const int nbox = BXY2 * (g.BZ + 2);
    const long nxy = (long)g.nx * g.ny;
    const S* x = (const S*)I.v;
    for (int u = threadIdx.x; u < 3 * nbox; u += kSweepThreads) {
        const int b = u / 3, q = u - 3 * b;
        const int i = TI * g.BX + b % BX2 - 1, j = TJ * g.BY + (b / BX2) % (g.BY + 2) - 1,
                  k = TK * g.BZ + b / BXY2 - 1;
        if (i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz)
            box[u] = ldCoh(x + 3 * ((long)i + (long)g.nx * j + nxy * k) + q);
    }
    __syncthreads();
    const int u = threadIdx.x;
    if (u >= 3 * g.BX * g.BY * g.BZ) return;
    const int c = u / 3, e = u - 3 * c;
    const int li = c % g.BX, lj = (c / g.BX) % g.BY, lk = c / (g.BX * g.BY);
    const int i = TI * g.BX + li, j = TJ * g.BY + lj, k = TK * g.BZ + lk;
    if (i >= g.nx || j >= g.ny || k >= g.nz) return;
    const long row = (long)i + (long)g.nx * ((long)j + (long)g.ny * k);
    const int pos = (li + 1) + BX2 * (lj + 1) + BXY2 * (lk + 1);
    const GS<int>* __restrict__ rp = toGlobal(fz.rp);
    const GS<int>* __restrict__ ci = toGlobal(fz.ci);
    const GS<S>* __restrict__ a = toGlobal((const S*)fz.a);
    const int k0 = rp[row], nb = rp[row + 1] - k0;
    S v[7][3];
    int off[7];
#pragma unroll
    for (int jb = 0; jb < 7; ++jb) {
        off[jb] = pos;
        if (jb < nb) {
            const long d = (long)ci[k0 + jb] - row;     // the 7-point neighbours only
            off[jb] = pos + (d == -nxy ? -BXY2 : d == -g.nx ? -BX2 : d == -1 ? -1 : d == 0 ? 0
                             : d == 1 ? 1 : d == g.nx ? BX2 : BXY2);
#pragma unroll
            for (int q = 0; q < 3; ++q) v[jb][q] = a[9L * (k0 + jb) + 3 * e + q];
        }
    }
    S acc = S(0);
#pragma unroll
    for (int jb = 0; jb < 7; ++jb) {
        if (jb < nb) {
            const S* xb = box + 3 * off[jb];
            S t = v[jb][0] * xb[0];
            t += v[jb][1] * xb[1];
            t += v[jb][2] * xb[2];
            acc += t;
        }
    }
    toGlobal((S*)fz.y)[3 * row + e] = acc;
}

// Forward sweep (Bwd = false): v = L^-1 d (unit lower), lower blocks in column
// order -z, -y, -x.  Backward sweep (Bwd = true): v = U^-1 v with the inverted
// (FP64) pivots, upper blocks in column order +x, +y, +z.
// Fused (backward, one instance): the ticket order also holds the operator's
// row blocks (inst < 0), each after every tile it reads (its own, the +x, +y,
// +z neighbours one backward level earlier and the -x, -y, -z neighbours one
// level later), and every tile publishes a flag once its results are stored.
template <class S, bool Bwd, bool Fused>
__global__ __launch_bounds__(kSweepThreads) void k_ilu_sweep_set(const IluInst* __restrict__ insts,
    const TileTask* __restrict__ order, int T, unsigned long long* counter, int* err, unsigned int* exits, int nInst,
    SpmvFuse fz)
{
    __shared__ S sV[(kPadMax + 1) * 3];       // the tile's cells and its ghost layer (padded layout),
                                              // then a zero triple read in place of absent neighbours
    __shared__ int sOrder[kStreamCells];      // tile positions in level order
    __shared__ int sDone[kMaxTileLevels];     // per cell level: compute waves that have finished it
    __shared__ int sWaves[kMaxTileLevels];    // per cell level: compute waves holding part of it
    __shared__ int sTicket;
    const int q = threadIdx.x & 3, e = q < 3 ? q : 2, pos = threadIdx.x >> 2;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (skipLaunch(insts, nInst, counter, exits)) return;
    // the epoch only moves when the last workgroup leaves, after every one has read it
    const int ep = Fused ? (int)__hip_atomic_load(fz.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
    int t = shareTicket(threadIdx.x == 0 ? drawTicket(counter, 0) : 0, &sTicket);
    for (;;) {
        if (t < 0 || t >= T) break;
        int next = 0;
        if (threadIdx.x == 0) next = drawTicket(counter, 0);
        const TileTask task = order[(Bwd && !Fused) ? T - 1 - t : t];
        if constexpr (Fused) {
            if (task.inst < 0) {
                spmvTile<S>(insts[0], fz, task.tile, ep, err, sV);
                __syncthreads();
                t = shareTicket(next, &sTicket);
                continue;
            }
        }
        const IluInst I = insts[task.inst];
        if (!isActive(I)) {
            t = shareTicket(next, &sTicket);
            continue;
        }
        const TileGeom& g = I.g;
        const int BX = g.BX, BY = g.BY, BZ = g.BZ, PX = BX + 1, PXY = (BX + 1) * (BY + 1);
        const int ncell = BX * BY * BZ;
        const long nxy = (long)g.nx * g.ny;
        int TI, TJ, TK;
        tileOf(g, task.tile, TI, TJ, TK);
        const int i0 = TI * BX, j0 = TJ * BY, k0 = TK * BZ;
        // every value slot starts empty; tile positions in level order
        for (int k = threadIdx.x; k < 3 * PXY * (BZ + 1); k += kSweepThreads) sV[k] = emptySlot<S>();
        if (threadIdx.x < 3) sV[3 * kPadMax + threadIdx.x] = S(0);
        if (threadIdx.x < BX + BY + BZ - 2 + 1) {
            const int l = threadIdx.x;
            sDone[l] = 0;
            sWaves[l] = (cellsBelow(l + 1, BX, BY, BZ) - 1) / 16 - cellsBelow(l, BX, BY, BZ) / 16 + 1;
        }
        if (threadIdx.x < ncell) {
            const int li = threadIdx.x % BX, lj = (threadIdx.x / BX) % BY, lk = threadIdx.x / (BX * BY);
            const int i = i0 + li, j = j0 + lj, k = k0 + lk;
            const int a = Bwd ? BX - 1 - li : li, b = Bwd ? BY - 1 - lj : lj, c = Bwd ? BZ - 1 - lk : lk;
            // dependency neighbours in column order: -z, -y, -x forward; +x, +y, +z backward
            const bool n0 = Bwd ? i + 1 < g.nx : k > 0;
            const bool n1 = Bwd ? j + 1 < g.ny : j > 0;
            const bool n2 = Bwd ? k + 1 < g.nz : i > 0;
            sOrder[cellsBelow(a + b + c, BX, BY, BZ) + rankInLevel(a, b, c, BX, BY)] =
                packCell(li, lj, lk, n0, n1, n2, i < g.nx && j < g.ny && k < g.nz);
        }
        __syncthreads();
        // exchange slots: direction d of cell I at xch[3 d N + 3 I + component]
        S* const xch = (S*)I.xch;
        const long N = nxy * g.nz;
        if (wave == kRecvWave) {
            // this lane's ghost values: gv = lane + 64 m runs over the tile's x-,
            // y- and z-face cells (three values each) whose neighbour exists
            const int nfx = BY * BZ, nfy = BX * BZ, nfz = BX * BY;
            const int nG = 3 * (nfx + nfy + nfz);
            // src: element offset of the slot in xch; dst: LDS index
            int src[kRecvPerLane], dst[kRecvPerLane];
            unsigned mask = 0;
#pragma clang loop unroll(full)
            for (int m = 0; m < kRecvPerLane; ++m) {
                const int gv = lane + 64 * m;
                src[m] = 0;
                dst[m] = 0;
                if (gv < nG) {
                    const int comp = gv % 3, f = gv / 3;
                    int dir, li, lj, lk;
                    if (f < nfx) {
                        dir = 0; lj = f % BY; lk = f / BY; li = Bwd ? BX - 1 : 0;
                    } else if (f < nfx + nfy) {
                        dir = 1; li = (f - nfx) % BX; lk = (f - nfx) / BX; lj = Bwd ? BY - 1 : 0;
                    } else {
                        dir = 2; li = (f - nfx - nfy) % BX; lj = (f - nfx - nfy) / BX; lk = Bwd ? BZ - 1 : 0;
                    }
                    const int i = i0 + li, j = j0 + lj, k = k0 + lk;
                    bool has = i < g.nx && j < g.ny && k < g.nz;
                    if (dir == 0) has = has && (Bwd ? i + 1 < g.nx : i > 0);
                    if (dir == 1) has = has && (Bwd ? j + 1 < g.ny : j > 0);
                    if (dir == 2) has = has && (Bwd ? k + 1 < g.nz : k > 0);
                    if (has) {
                        const long Ic = (long)i + (long)g.nx * ((long)j + (long)g.ny * k);
                        const long stride = dir == 0 ? 1 : (dir == 1 ? (long)g.nx : nxy);
                        const int pTile = Bwd ? li + PX * lj + PXY * lk : (li + 1) + PX * (lj + 1) + PXY * (lk + 1);
                        const int pStride = dir == 0 ? 1 : (dir == 1 ? PX : PXY);
                        src[m] = (int)(3 * dir * N + 3 * (Bwd ? Ic + stride : Ic - stride) + comp);
                        dst[m] = 3 * (Bwd ? pTile + pStride : pTile - pStride) + comp;
                        mask |= 1u << m;
                    }
                }
            }
            // continuous polling: every slot always has kRecvDepth loads in flight;
            // a load is re-issued as soon as its result has been looked at, so an
            // arriving value is seen about 1/kRecvDepth of a memory latency (plus
            // the return trip) after it lands.  All loads are
            // unconditional (finished or unused slots re-read a harmless address) and
            // the loop has no stores, so each wait covers exactly the one load it needs.
            const unsigned all = mask;
            S val[kRecvDepth][kRecvPerLane];
#pragma clang loop unroll(full)
            for (int d = 0; d < kRecvDepth; ++d)
#pragma clang loop unroll(full)
                for (int m = 0; m < kRecvPerLane; ++m) val[d][m] = ldCoh(xch + src[m]);
            int failed = loadRelaxed(err);
            long spins = 0;
            while (mask) {
#pragma clang loop unroll(full)
                for (int d = 0; d < kRecvDepth; ++d)
#pragma clang loop unroll(full)
                    for (int m = 0; m < kRecvPerLane; ++m) {
                        const S x = val[d][m];
                        val[d][m] = ldCoh(xch + src[m]);
                        if (((mask >> m) & 1u) && !isEmpty(x)) {
                            ldsStore(sV + dst[m], x);
                            mask &= ~(1u << m);
                        }
                    }
                const int f = failed;
                failed = loadRelaxed(err);
                if (f || ++spins > kSpinLimit) {
                    if (!f) atomicOr(err, 1);
                    break;
                }
            }
            // empty the slots taken (the next sweep's values will arrive in them)
#pragma clang loop unroll(full)
            for (int m = 0; m < kRecvPerLane; ++m)
                if ((all >> m) & 1u) stCoh(xch + src[m], emptySlot<S>());
        } else if (16 * wave < ncell) {
            const int word = pos < ncell ? sOrder[pos] : 0;
            const bool valid = pos < ncell && (word & (1 << 30));
            const int li = word & 0xFF, lj = (word >> 8) & 0xFF, lk = (word >> 16) & 0xFF;
            const bool h0 = valid && (word & (1 << 27)), h1 = valid && (word & (1 << 28));
            const bool h2 = valid && (word & (1 << 29));
            const long Ic = (long)(i0 + li) + (long)g.nx * ((long)(j0 + lj) + (long)g.ny * (k0 + lk));
            const int p = Bwd ? li + PX * lj + PXY * lk : (li + 1) + PX * (lj + 1) + PXY * (lk + 1);
            // dependency positions; an absent neighbour reads the zero triple
            const int d0 = h0 ? (Bwd ? p + 1 : p - PXY) : kPadMax;
            const int d1 = h1 ? (Bwd ? p + PX : p - PX) : kPadMax;
            const int d2 = h2 ? (Bwd ? p + PXY : p - 1) : kPadMax;
            auto levelOf = [&](int w) {
                const int a = w & 0xFF, b = (w >> 8) & 0xFF, c = (w >> 16) & 0xFF;
                return Bwd ? (BX - 1 - a) + (BY - 1 - b) + (BZ - 1 - c) : a + b + c;
            };
            const GS<S>* __restrict__ lu = toGlobal((const S*)I.lu);
            GS<S>* __restrict__ v = toGlobal((S*)I.v);
            // this lane's row of the cell's off-diagonal blocks (and pivot) and its right-hand side
            S B0[3] = {}, B1[3] = {}, B2[3] = {};
            double D[3] = {};
            S r = S(0);
            if (valid) {
                long o;
                if (Bwd) {
                    r = v[3 * Ic + e];
                    o = toGlobal(I.diag)[Ic] + 1;
                    for (int w = 0; w < 3; ++w) D[w] = toGlobal(I.dinv)[9 * Ic + 3 * e + w];
                } else {
                    r = toGlobal((const S*)I.d)[3 * Ic + e];
                    o = toGlobal(I.rp)[Ic];
                }
                if (h0) { for (int w = 0; w < 3; ++w) B0[w] = lu[9 * o + 3 * e + w]; ++o; }
                if (h1) { for (int w = 0; w < 3; ++w) B1[w] = lu[9 * o + 3 * e + w]; ++o; }
                if (h2) { for (int w = 0; w < 3; ++w) B2[w] = lu[9 * o + 3 * e + w]; }
            }
            // the faces this lane's row is sent across
            const int i = i0 + li, j = j0 + lj, k = k0 + lk;
            const bool row = valid && q < 3;
            const bool s0 = row && (Bwd ? li == 0 && i > 0 : li + 1 == BX && i + 1 < g.nx);
            const bool s1 = row && (Bwd ? lj == 0 && j > 0 : lj + 1 == BY && j + 1 < g.ny);
            const bool s2 = row && (Bwd ? lk == 0 && k > 0 : lk + 1 == BZ && k + 1 < g.nz);
            S* const own = xch + 3 * Ic + e;
            // the rows are in registers before the levels start: the loop below then
            // has only stores in flight, which none of its waits has to cover
            waitStores();
            const int myLev = valid ? levelOf(word) : -1;
            const int lastLev = levelOf(sOrder[min(16 * wave + 15, ncell - 1)]);
            S res = S(0);
            const int firstLev = levelOf(sOrder[16 * wave]);
            for (int lev = firstLev; lev <= lastLev; ++lev) {
                // before its first level a wave waits until level lev-2 is complete
                // (every wave holding part of it has counted it; one broadcast LDS
                // word per poll), so the waves further ahead do not load the LDS with
                // value polls.  The count only throttles: the value polls below carry
                // the dependencies, and a wave's later levels follow its own results
                // without it.
                if (lev > 1 && lev == firstLev) {
                    const int waves = sWaves[lev - 2];
                    int spins = 0;
                    while (ldsLoad(sDone + lev - 2) < waves) {
                        if (++spins > kSpinLimit || ((spins & 255) == 0 && loadRelaxed(err))) {
                            atomicOr(err, 1);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                if (myLev == lev) {
                    // the dependencies themselves (level lev-1 and ghost values) are
                    // polled: a slot leaves the sentinel exactly once per tile, with its
                    // final value, so the last step of the chain costs no count
                    // (the wave polls until all its lanes of the level are ready, so the
                    // loop and its spin count are wave-uniform)
                    S x0[3], x1[3], x2[3];
                    int spins = 0;
                    for (;;) {
                        for (int w = 0; w < 3; ++w) {
                            x0[w] = ldsLoad(sV + 3 * d0 + w);
                            x1[w] = ldsLoad(sV + 3 * d1 + w);
                            x2[w] = ldsLoad(sV + 3 * d2 + w);
                        }
                        bool empty = false;
                        for (int w = 0; w < 3; ++w) empty |= isEmpty(x0[w]) | isEmpty(x1[w]) | isEmpty(x2[w]);
                        if (!__any(empty)) break;
                        if (++spins > kSpinLimit || ((spins & 255) == 0 && loadRelaxed(err))) {
                            atomicOr(err, 1);
                            break;
                        }
                    }
                    const S t0 = rowDot(B0, x0), t1 = rowDot(B1, x1), t2 = rowDot(B2, x2);
                    if (h0) r += S(-1) * t0;
                    if (h1) r += S(-1) * t1;
                    if (h2) r += S(-1) * t2;
                    S out = r;
                    if (Bwd) {
                        const S r0 = quadLane<0>(r), r1 = quadLane<1>(r), r2 = quadLane<2>(r);
                        double w = D[0] * (double)r0;
                        w += D[1] * (double)r1;
                        w += D[2] * (double)r2;
                        out = S(0) + S(1) * (S)w;
                    }
                    res = out;
                    if (q < 3) {
                        ldsStore(sV + 3 * p + e, isEmpty(out) ? S(NAN) : out);
                        if (s0) sendSlot(own, out);
                        if (s1) sendSlot(own + 3 * N, out);
                        if (s2) sendSlot(own + 6 * N, out);
                    }
                }
                // count the level (a throttle only, so no wait for the results first)
                if (lane == 0) __hip_atomic_fetch_add(sDone + lev, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            // the tile's own results leave after its levels, behind the face values
            // the neighbour tiles are waiting for (device-scope when other
            // workgroups of this launch read them)
            if (row) {
                if constexpr (Fused) stCoh((S*)I.v + 3 * Ic + e, res);
                else v[3 * Ic + e] = res;
            }
        }
        if constexpr (Fused) {
            waitStores();
            __syncthreads();
            if (threadIdx.x == 0) stCoh(fz.tileDone + task.tile, ep);
        }
        __syncthreads();
        t = shareTicket(next, &sTicket);
    }
    if constexpr (Fused) {
        if (threadIdx.x == 0 && atomicAdd(exits, 1u) == gridDim.x - 1) {
            atomicExch(counter, 0ULL);
            atomicAdd(fz.epoch, 1u);
            atomicExch(exits, 0u);
        }
    } else {
        leaveLaunch(counter, exits);
    }
}

// Batched sweeps: the same lane layout without a receiving wave.  A tile waits
// for all of its neighbour tiles' face values, stages them as the LDS ghost
// layer, runs its levels separated by workgroup barriers and sends its own
// face values after its last level.  A level costs less than in the streaming
// sweeps (one barrier instead of per-level counts), a tile hop more (a tile
// starts when its upstream tiles have finished); with many more tiles per tile
// level than resident workgroups (large ensembles) the sweeps are throughput-
// rather than latency-bound and this form is the faster one.
// Forward sweep: v = L^-1 d (unit lower), lower blocks in column order -z,-y,-x.
template <class S>
__global__ __launch_bounds__(kSweepThreads) void k_ilu_fwd_batch(const IluInst* __restrict__ insts,
    const TileTask* __restrict__ order, int T, unsigned long long* counter, int* err, unsigned int* exits, int nInst)
{
    __shared__ S sV[kPadMax * 3];          // results of the tile's cells (padded layout)
    __shared__ int sOrder[kStreamCells];   // cells in level order
    __shared__ int sTicket;
    const int q = threadIdx.x & 3, e = q < 3 ? q : 2, pos = threadIdx.x >> 2;
    if (skipLaunch(insts, nInst, counter, exits)) return;
    int t = shareTicket(threadIdx.x == 0 ? drawTicket(counter, 0) : 0, &sTicket);
    for (;;) {
        if (t < 0 || t >= T) break;
        int next = 0;
        if (threadIdx.x == 0) next = drawTicket(counter, 0);
        const TileTask task = order[t];
        const IluInst I = insts[task.inst];
        if (!isActive(I)) {
            t = shareTicket(next, &sTicket);
            continue;
        }
        const TileGeom& g = I.g;
        const int BX = g.BX, BY = g.BY, BZ = g.BZ, PX = BX + 1, PXY = (BX + 1) * (BY + 1);
        const GS<S>* __restrict__ lu = toGlobal((const S*)I.lu);
        const GS<S>* __restrict__ d = toGlobal((const S*)I.d);
        GS<S>* __restrict__ v = toGlobal((S*)I.v);
        const long nxy = (long)g.nx * g.ny;
        int TI, TJ, TK;
        tileOf(g, task.tile, TI, TJ, TK);
        if (threadIdx.x < BX * BY * BZ) {
            const int li = threadIdx.x % BX, lj = (threadIdx.x / BX) % BY, lk = threadIdx.x / (BX * BY);
            const int i = TI * BX + li, j = TJ * BY + lj, k = TK * BZ + lk;
            sOrder[cellsBelow(li + lj + lk, BX, BY, BZ) + rankInLevel(li, lj, lk, BX, BY)] =
                (i < g.nx && j < g.ny && k < g.nz) ? packCell(li, lj, lk, k > 0, j > 0, i > 0, true) : -1;
        }
        __syncthreads();
        const int word = pos < BX * BY * BZ ? sOrder[pos] : -1;
        const int li = word & 0xFF, lj = (word >> 8) & 0xFF, lk = (word >> 16) & 0xFF;
        const bool hz = word >= 0 && (word & (1 << 27)), hy = word >= 0 && (word & (1 << 28));
        const bool hx = word >= 0 && (word & (1 << 29));
        const long Ic = (long)(TI * BX + li) + (long)g.nx * ((long)(TJ * BY + lj) + (long)g.ny * (TK * BZ + lk));
        const int p = (li + 1) + PX * (lj + 1) + PXY * (lk + 1);
        const int myLev = word >= 0 ? li + lj + lk : -1;
        // this lane's row of the cell's lower blocks and its right-hand side
        S Lz[3] = {}, Ly[3] = {}, Lx[3] = {};
        S r = S(0);
        if (word >= 0) {
            r = d[3 * Ic + e];
            long o = toGlobal(I.rp)[Ic];
            if (hz) { for (int w = 0; w < 3; ++w) Lz[w] = lu[9 * o + 3 * e + w]; ++o; }
            if (hy) { for (int w = 0; w < 3; ++w) Ly[w] = lu[9 * o + 3 * e + w]; ++o; }
            if (hx) { for (int w = 0; w < 3; ++w) Lx[w] = lu[9 * o + 3 * e + w]; }
        }
        // ghosts: the neighbour tiles' face values, received through the exchange
        // slots in one batch before the tile's levels (receiving them level by
        // level would put a memory round trip on every level)
        S* const xx = (S*)I.xch;
        S* const xy = xx + 3 * nxy * g.nz;
        S* const xz = xx + 6 * nxy * g.nz;
        if (q < 3) {
            if (hz && lk == 0) sV[3 * (p - PXY) + e] = receiveSlot(xz + 3 * (Ic - nxy) + e, err);
            if (hy && lj == 0) sV[3 * (p - PX) + e] = receiveSlot(xy + 3 * (Ic - g.nx) + e, err);
            if (hx && li == 0) sV[3 * (p - 1) + e] = receiveSlot(xx + 3 * (Ic - 1) + e, err);
        }
        __syncthreads();
        const int nLev = BX + BY + BZ - 2;
        for (int lev = 0; lev <= nLev; ++lev) {
            if (myLev == lev) {
                const S sz = rowDot(Lz, sV + 3 * (p - PXY));
                const S sy = rowDot(Ly, sV + 3 * (p - PX));
                const S sx = rowDot(Lx, sV + 3 * (p - 1));
                if (hz) r += S(-1) * sz;
                if (hy) r += S(-1) * sy;
                if (hx) r += S(-1) * sx;
                if (q < 3) sV[3 * p + e] = r;
            }
            __syncthreads();
        }
        if (word >= 0 && q < 3) {
            v[3 * Ic + e] = r;
            if (li + 1 == BX && TI * BX + li + 1 < g.nx) sendSlot(xx + 3 * Ic + e, r);
            if (lj + 1 == BY && TJ * BY + lj + 1 < g.ny) sendSlot(xy + 3 * Ic + e, r);
            if (lk + 1 == BZ && TK * BZ + lk + 1 < g.nz) sendSlot(xz + 3 * Ic + e, r);
        }
        __syncthreads();
        t = shareTicket(next, &sTicket);
    }
    leaveLaunch(counter, exits);
}

// Backward sweep: v = U^-1 v with the inverted (FP64) pivots; upper blocks +x,+y,+z.
template <class S>
__global__ __launch_bounds__(kSweepThreads) void k_ilu_bwd_batch(const IluInst* __restrict__ insts,
    const TileTask* __restrict__ order, int T, unsigned long long* counter, int* err, unsigned int* exits, int nInst)
{
    __shared__ S sV[kPadMax * 3];          // results of the tile's cells (padded layout)
    __shared__ int sOrder[kStreamCells];
    __shared__ int sTicket;
    const int q = threadIdx.x & 3, e = q < 3 ? q : 2, pos = threadIdx.x >> 2;
    if (skipLaunch(insts, nInst, counter, exits)) return;
    int t = shareTicket(threadIdx.x == 0 ? drawTicket(counter, 0) : 0, &sTicket);
    for (;;) {
        if (t < 0 || t >= T) break;
        int next = 0;
        if (threadIdx.x == 0) next = drawTicket(counter, 0);
        const TileTask task = order[T - 1 - t];
        const IluInst I = insts[task.inst];
        if (!isActive(I)) {
            t = shareTicket(next, &sTicket);
            continue;
        }
        const TileGeom& g = I.g;
        const int BX = g.BX, BY = g.BY, BZ = g.BZ, PX = BX + 1, PXY = (BX + 1) * (BY + 1);
        const GS<S>* __restrict__ lu = toGlobal((const S*)I.lu);
        const GS<double>* __restrict__ dinv = toGlobal(I.dinv);
        GS<S>* __restrict__ v = toGlobal((S*)I.v);
        const long nxy = (long)g.nx * g.ny;
        int TI, TJ, TK;
        tileOf(g, task.tile, TI, TJ, TK);
        if (threadIdx.x < BX * BY * BZ) {
            const int li = threadIdx.x % BX, lj = (threadIdx.x / BX) % BY, lk = threadIdx.x / (BX * BY);
            const int i = TI * BX + li, j = TJ * BY + lj, k = TK * BZ + lk;
            const int a = BX - 1 - li, b = BY - 1 - lj, c = BZ - 1 - lk;
            sOrder[cellsBelow(a + b + c, BX, BY, BZ) + rankInLevel(a, b, c, BX, BY)] =
                (i < g.nx && j < g.ny && k < g.nz)
                    ? packCell(li, lj, lk, i + 1 < g.nx, j + 1 < g.ny, k + 1 < g.nz, true) : -1;
        }
        __syncthreads();
        const int word = pos < BX * BY * BZ ? sOrder[pos] : -1;
        const int li = word & 0xFF, lj = (word >> 8) & 0xFF, lk = (word >> 16) & 0xFF;
        const bool ux = word >= 0 && (word & (1 << 27)), uy = word >= 0 && (word & (1 << 28));
        const bool uz = word >= 0 && (word & (1 << 29));
        const long Ic = (long)(TI * BX + li) + (long)g.nx * ((long)(TJ * BY + lj) + (long)g.ny * (TK * BZ + lk));
        const int p = li + PX * lj + PXY * lk;
        const int myLev = word >= 0 ? (BX - 1 - li) + (BY - 1 - lj) + (BZ - 1 - lk) : -1;
        S Ux[3] = {}, Uy[3] = {}, Uz[3] = {};
        double D[3] = {};
        S r = S(0);
        if (word >= 0) {
            r = v[3 * Ic + e];
            long o = toGlobal(I.diag)[Ic] + 1;
            if (ux) { for (int w = 0; w < 3; ++w) Ux[w] = lu[9 * o + 3 * e + w]; ++o; }
            if (uy) { for (int w = 0; w < 3; ++w) Uy[w] = lu[9 * o + 3 * e + w]; ++o; }
            if (uz) { for (int w = 0; w < 3; ++w) Uz[w] = lu[9 * o + 3 * e + w]; }
            for (int w = 0; w < 3; ++w) D[w] = dinv[9 * Ic + 3 * e + w];
        }
        S* const xx = (S*)I.xch;
        S* const xy = xx + 3 * nxy * g.nz;
        S* const xz = xx + 6 * nxy * g.nz;
        if (q < 3) {
            if (ux && li + 1 == BX) sV[3 * (p + 1) + e] = receiveSlot(xx + 3 * (Ic + 1) + e, err);
            if (uy && lj + 1 == BY) sV[3 * (p + PX) + e] = receiveSlot(xy + 3 * (Ic + g.nx) + e, err);
            if (uz && lk + 1 == BZ) sV[3 * (p + PXY) + e] = receiveSlot(xz + 3 * (Ic + nxy) + e, err);
        }
        __syncthreads();
        const int nLev = BX + BY + BZ - 2;
        S out = S(0);
        for (int lev = 0; lev <= nLev; ++lev) {
            if (myLev == lev) {
                const S sx = rowDot(Ux, sV + 3 * (p + 1));
                const S sy = rowDot(Uy, sV + 3 * (p + PX));
                const S sz = rowDot(Uz, sV + 3 * (p + PXY));
                if (ux) r += S(-1) * sx;
                if (uy) r += S(-1) * sy;
                if (uz) r += S(-1) * sz;
                const S r0 = quadLane<0>(r), r1 = quadLane<1>(r), r2 = quadLane<2>(r);
                double w = D[0] * (double)r0;
                w += D[1] * (double)r1;
                w += D[2] * (double)r2;
                out = S(0) + S(1) * (S)w;
                if (q < 3) sV[3 * p + e] = out;
            }
            __syncthreads();
        }
        if (word >= 0 && q < 3) {
            v[3 * Ic + e] = out;
            if (li == 0 && TI * BX + li > 0) sendSlot(xx + 3 * Ic + e, out);
            if (lj == 0 && TJ * BY + lj > 0) sendSlot(xy + 3 * Ic + e, out);
            if (lk == 0 && TK * BZ + lk > 0) sendSlot(xz + 3 * Ic + e, out);
        }
        __syncthreads();
        t = shareTicket(next, &sTicket);
    }
    leaveLaunch(counter, exits);
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

// One workgroup of kSweepThreads per tile with the sweeps' lane layout: four
// lanes per cell, cells in level order, lane e (the fourth repeats row 2)
// forming row e of every block product of its cell.  The inverted pivots of
// the tile's cells and of the neighbour tiles' face cells sit in LDS (padded
// layout), as do the cells' U_JI blocks; the cell's pivot row and lower-block
// rows stay in registers.  After the pivot update each lane gathers the whole
// pivot block from its quad (DPP) and inverts it with inv3, so every entry
// is computed with the operations, and in the order, of the one-thread-per-
// cell form.  Levels are separated by workgroup barriers; a tile waits for
// the done flags of its -x, -y, -z neighbour tiles, publishes its own once
// its inverted pivots are acknowledged, and stores its factor blocks after.
__global__ __launch_bounds__(kSweepThreads) void k_ilu_factor_set(const IluInst* __restrict__ insts,
    const TileTask* __restrict__ order, int T, unsigned long long* counter, long long base, int epoch, int* err,
    int* fail, unsigned int* exits)
{
    __shared__ double sP[kPadMax * 9];            // inverted pivots (padded tile + ghost layer)
    __shared__ double sU[kStreamCells * 27];      // per cell (level-order rank): U_JI of its -z, -y, -x neighbours
    __shared__ int sOrder[kStreamCells];
    __shared__ int sTicket;
    const int q = threadIdx.x & 3, e = q < 3 ? q : 2, pos = threadIdx.x >> 2;
    int t = shareTicket(threadIdx.x == 0 ? drawTicket(counter, base) : 0, &sTicket);
    for (;;) {
        if (t < 0 || t >= T) break;
        int next = 0;
        if (threadIdx.x == 0) next = drawTicket(counter, base);
        const TileTask task = order[t];
        const IluInst I = insts[task.inst];
        if (!isActive(I)) {
            publish(I.done, task.tile, epoch);
            t = shareTicket(next, &sTicket);
            continue;
        }
        const TileGeom& g = I.g;
        const int BX = g.BX, BY = g.BY, BZ = g.BZ, PX = BX + 1, PXY = (BX + 1) * (BY + 1);
        const int ncell = BX * BY * BZ;
        const long nxy = (long)g.nx * g.ny;
        int TI, TJ, TK;
        tileOf(g, task.tile, TI, TJ, TK);
        const int i0 = TI * BX, j0 = TJ * BY, k0 = TK * BZ;
        if (threadIdx.x < ncell) {
            const int li = threadIdx.x % BX, lj = (threadIdx.x / BX) % BY, lk = threadIdx.x / (BX * BY);
            const int i = i0 + li, j = j0 + lj, k = k0 + lk;
            sOrder[cellsBelow(li + lj + lk, BX, BY, BZ) + rankInLevel(li, lj, lk, BX, BY)] =
                packCell(li, lj, lk, k > 0, j > 0, i > 0, i < g.nx && j < g.ny && k < g.nz);
        }
        __syncthreads();
        const int word = pos < ncell ? sOrder[pos] : 0;
        const bool valid = pos < ncell && (word & (1 << 30));
        const int li = word & 0xFF, lj = (word >> 8) & 0xFF, lk = (word >> 16) & 0xFF;
        const bool hz = valid && (word & (1 << 27)), hy = valid && (word & (1 << 28));
        const bool hx = valid && (word & (1 << 29));
        const int i = i0 + li, j = j0 + lj, k = k0 + lk;
        const long Ic = (long)i + (long)g.nx * ((long)j + (long)g.ny * k);
        const int p = (li + 1) + PX * (lj + 1) + PXY * (lk + 1);
        const GS<int>* __restrict__ rp = toGlobal(I.rp);
        const GS<int>* __restrict__ dg = toGlobal(I.diag);
        double* __restrict__ lu = (double*)I.luw;
        double* __restrict__ dinv = I.dinvw;
        // this lane's row of the cell's pivot and lower blocks; its row of the
        // U_JI blocks goes to LDS (every lane of the cell reads the whole block)
        double D[3] = {}, Az[3] = {}, Ay[3] = {}, Ax[3] = {};
        long oz = 0, oy = 0, ox = 0, od = 0;
        if (valid) {
            long o = rp[Ic];
            if (hz) oz = o++;
            if (hy) oy = o++;
            if (hx) ox = o++;
            od = dg[Ic];
            const GS<double>* __restrict__ a = toGlobal((const double*)lu);
            for (int w = 0; w < 3; ++w) D[w] = a[9 * od + 3 * e + w];
            if (hz) {
                const long oJ = dg[Ic - nxy] + 1 + (i + 1 < g.nx) + (j + 1 < g.ny);
                for (int w = 0; w < 3; ++w) { Az[w] = a[9 * oz + 3 * e + w]; if (q < 3) sU[27 * pos + 3 * e + w] = a[9 * oJ + 3 * e + w]; }
            }
            if (hy) {
                const long oJ = dg[Ic - g.nx] + 1 + (i + 1 < g.nx);
                for (int w = 0; w < 3; ++w) { Ay[w] = a[9 * oy + 3 * e + w]; if (q < 3) sU[27 * pos + 9 + 3 * e + w] = a[9 * oJ + 3 * e + w]; }
            }
            if (hx) {
                const long oJ = dg[Ic - 1] + 1;
                for (int w = 0; w < 3; ++w) { Ax[w] = a[9 * ox + 3 * e + w]; if (q < 3) sU[27 * pos + 18 + 3 * e + w] = a[9 * oJ + 3 * e + w]; }
            }
        }
        waitDeps(g, TI, TJ, TK, -1, I.done, epoch, err);
        // inverted pivots of the neighbours across the tile faces, one round trip
        if (q < 3) {
            if (hz && lk == 0) for (int w = 0; w < 3; ++w) sP[9 * (p - PXY) + 3 * e + w] = ldCoh(dinv + 9 * (Ic - nxy) + 3 * e + w);
            if (hy && lj == 0) for (int w = 0; w < 3; ++w) sP[9 * (p - PX) + 3 * e + w] = ldCoh(dinv + 9 * (Ic - g.nx) + 3 * e + w);
            if (hx && li == 0) for (int w = 0; w < 3; ++w) sP[9 * (p - 1) + 3 * e + w] = ldCoh(dinv + 9 * (Ic - 1) + 3 * e + w);
        }
        __syncthreads();
        const int myLev = valid ? li + lj + lk : -1;
        const int nLev = BX + BY + BZ - 2;
        for (int lev = 0; lev <= nLev; ++lev) {
            if (myLev == lev) {
                // L_IJ = A_IJ Dj^-1 (row e), then A_II -= L_IJ A_JI, for J = -z, -y, -x
                auto update = [&](double* Arow, int pn, int ub) {
                    double L[3], Tm[3];
                    const double* Dj = sP + 9 * pn;
                    const double* U = sU + 27 * pos + ub;
                    for (int c2 = 0; c2 < 3; ++c2) {
                        double s2 = Arow[0] * Dj[c2];
                        s2 += Arow[1] * Dj[3 + c2];
                        s2 += Arow[2] * Dj[6 + c2];
                        L[c2] = s2;
                    }
                    for (int c2 = 0; c2 < 3; ++c2) Arow[c2] = L[c2];
                    for (int c2 = 0; c2 < 3; ++c2) {
                        double s2 = L[0] * U[c2];
                        s2 += L[1] * U[3 + c2];
                        s2 += L[2] * U[6 + c2];
                        Tm[c2] = s2;
                    }
                    for (int c2 = 0; c2 < 3; ++c2) D[c2] -= Tm[c2];
                };
                if (hz) update(Az, p - PXY, 0);
                if (hy) update(Ay, p - PX, 9);
                if (hx) update(Ax, p - 1, 18);
                // the whole pivot block from the quad's three rows, inverted as one
                double M[9];
                for (int c2 = 0; c2 < 3; ++c2) {
                    M[c2] = quadLane<0>(D[c2]);
                    M[3 + c2] = quadLane<1>(D[c2]);
                    M[6 + c2] = quadLane<2>(D[c2]);
                }
                if (!inv3(M) && q == 0) atomicOr(fail, 2);
                for (int c2 = 0; c2 < 3; ++c2) D[c2] = M[3 * e + c2];
                if (q < 3) for (int c2 = 0; c2 < 3; ++c2) sP[9 * p + 3 * e + c2] = D[c2];
            }
            __syncthreads();
        }
        // inverted pivots first (dependent tiles read them), then the flag, then
        // the factor blocks, read only by later kernels
        const bool row = valid && q < 3;
        if (row) for (int c2 = 0; c2 < 3; ++c2) stCoh(dinv + 9 * Ic + 3 * e + c2, D[c2]);
        publish(I.done, task.tile, epoch);
        if (row) {
            GS<double>* __restrict__ w = toGlobal(lu);
            if (hz) for (int c2 = 0; c2 < 3; ++c2) w[9 * oz + 3 * e + c2] = Az[c2];
            if (hy) for (int c2 = 0; c2 < 3; ++c2) w[9 * oy + 3 * e + c2] = Ay[c2];
            if (hx) for (int c2 = 0; c2 < 3; ++c2) w[9 * ox + 3 * e + c2] = Ax[c2];
            for (int c2 = 0; c2 < 3; ++c2) w[9 * od + 3 * e + c2] = D[c2];
        }
        t = shareTicket(next, &sTicket);
    }
    leaveLaunch(counter, exits);
}

// ---------------------------------------------------------------------------
// empty exchange slots
__global__ void k_fill_empty64(unsigned long long* p, long n)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = kEmpty64;
}
__global__ void k_fill_empty32(unsigned int* p, long n)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = kEmpty32;
}

namespace {
// what the sweep kernels can hold: cells per tile, the tile and its ghost
// layer in LDS, and the ghost values of the receiving wave
bool tileFits(int bx, int by, int bz)
{
    return bx * by * bz <= kStreamCells && (bx + 1) * (by + 1) * (bz + 1) <= kPadMax &&
           3 * (by * bz + bx * bz + bx * by) <= 64 * kRecvPerLane && bx + by + bz - 2 < kMaxTileLevels;
}
}  // namespace

TileGeom makeTileGeom(int nx, int ny, int nz, int bx, int by, int bz)
{
    if (bx <= 0 || by <= 0 || bz <= 0) return makeTileGeom(nx, ny, nz);
    if (!tileFits(bx, by, bz)) throw std::invalid_argument("tile shape exceeds the sweep kernels' limits");
    TileGeom g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.BX = bx; g.BY = by; g.BZ = bz;
    g.TX = (nx + bx - 1) / bx; g.TY = (ny + by - 1) / by; g.TZ = (nz + bz - 1) / bz;
    g.T = g.TX * g.TY * g.TZ;
    return g;
}

std::vector<TileGeom> tileShapeCandidates(int nx, int ny, int nz)
{
    static const int shapes[][3] = {{6, 8, 5}, {8, 6, 5}, {8, 8, 3}, {10, 6, 4}, {6, 10, 4}, {12, 5, 4},
                                    {6, 6, 6}, {16, 5, 3}, {5, 16, 3}, {4, 6, 10}, {12, 12, 1}, {16, 8, 1}};
    std::vector<TileGeom> out;
    out.push_back(makeTileGeom(nx, ny, nz));
    for (const auto& sh : shapes) {
        // no tile much larger than the grid in any direction
        if (sh[0] > 2 * nx || sh[1] > 2 * ny || sh[2] > 2 * nz) continue;
        const int bx = std::min(sh[0], nx), by = std::min(sh[1], ny), bz = std::min(sh[2], nz);
        if (!tileFits(bx, by, bz)) continue;
        const TileGeom g = makeTileGeom(nx, ny, nz, bx, by, bz);
        bool dup = false;
        for (const auto& o : out) dup = dup || (o.BX == g.BX && o.BY == g.BY && o.BZ == g.BZ);
        if (!dup) out.push_back(g);
    }
    return out;
}

TileGeom makeTileGeom(int nx, int ny, int nz)
{
    // tile shape: about kStreamCells cells, as cubic as the grid allows
    const int bz = std::min(nz, 5);
    const int rest = kStreamCells / bz;
    int bx = std::min(nx, std::max(1, (int)std::sqrt((double)rest)));
    int by = std::min(ny, rest / bx);
    bx = std::min(nx, rest / by);
    int cz = bz;
    while (!tileFits(bx, by, cz)) {
        if (by >= bx && by >= cz) --by;
        else if (bx >= cz) --bx;
        else --cz;
    }
    return makeTileGeom(nx, ny, nz, bx, by, cz);
}

TiledIluSet::TiledIluSet(hipStream_t s)
{
    counter_.alloc(1);
    OM_HIP(hipMemsetAsync(counter_.p, 0, sizeof(unsigned long long), s));
    exits_.alloc(1);
    OM_HIP(hipMemsetAsync(exits_.p, 0, sizeof(unsigned int), s));
    err_.alloc(1);
    OM_HIP(hipMemsetAsync(err_.p, 0, sizeof(int), s));
    int dev = 0;
    OM_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    OM_HIP(hipGetDeviceProperties(&prop, dev));
    // workgroups of the factorization and of the sweeps that fit on the device
    // at once (the factorization holds far more registers per thread); the
    // launches use no more, so no workgroup waits for another one to leave
    auto resident = [&](std::initializer_list<const void*> ks, int threads) {
        int perCU = 1 << 20;
        for (const void* k : ks) {
            int n = 0;
            OM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, threads, 0));
            perCU = std::min(perCU, n);
        }
        if (perCU < 1) throw std::runtime_error("tiled ILU kernels cannot be resident");
        return perCU * prop.multiProcessorCount;
    };
    residentFactor_ = resident({(const void*)k_ilu_factor_set}, kSweepThreads);
    residentSweep_[0][0] = resident({(const void*)k_ilu_sweep_set<double, false, false>}, kSweepThreads);
    residentSweep_[0][1] = resident({(const void*)k_ilu_sweep_set<double, true, false>}, kSweepThreads);
    residentSweep_[1][0] = resident({(const void*)k_ilu_sweep_set<float, false, false>}, kSweepThreads);
    residentSweep_[1][1] = resident({(const void*)k_ilu_sweep_set<float, true, false>}, kSweepThreads);
    residentFused_[0] = resident({(const void*)k_ilu_sweep_set<double, true, true>}, kSweepThreads);
    residentFused_[1] = resident({(const void*)k_ilu_sweep_set<float, true, true>}, kSweepThreads);
    residentBatch_[0][0] = resident({(const void*)k_ilu_fwd_batch<double>}, kSweepThreads);
    residentBatch_[0][1] = resident({(const void*)k_ilu_bwd_batch<double>}, kSweepThreads);
    residentBatch_[1][0] = resident({(const void*)k_ilu_fwd_batch<float>}, kSweepThreads);
    residentBatch_[1][1] = resident({(const void*)k_ilu_bwd_batch<float>}, kSweepThreads);
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
    cellOff_.assign(geoms.size() + 1, 0);
    for (size_t b = 0; b < geoms.size(); ++b) {
        const long n = (long)geoms[b].nx * geoms[b].ny * geoms[b].nz;
        // the sweeps address an instance's exchange slots with 32-bit offsets
        if (9 * n >= (1L << 31)) throw std::invalid_argument("tiled ILU: more than 238 million cells in one grid");
        cellOff_[b + 1] = cellOff_[b] + n;
    }
    OM_HIP(hipStreamSynchronize(s));
    T_ = (int)ord.size();
    order_.upload(ord.data(), ord.size(), s);
    done_.alloc(std::max(T_, 1));
    OM_HIP(hipMemsetAsync(done_.p, 0, std::max(T_, 1) * sizeof(int), s));
    // flags restart from zero: restart the epoch so that no stale value matches
    epoch_ = 0;
    // exchange slots of the sweeps: FP64 now, FP32 on first use
    xch64_.alloc(std::max<long>(9 * cellOff_.back(), 1));
    xch32_.release();
    resetExchange(s);
    gridFactor_ = std::min(T_, residentFactor_);
    // streaming pays off while a tile level has fewer tiles than there are
    // resident workgroups (the sweeps wait on the dependency chain); with more,
    // they are throughput-bound and the batched form does less work per tile
    int levels = 1;
    for (const TileGeom& g : geoms) levels = std::max(levels, g.TX + g.TY + g.TZ - 2);
    autoStream_ = (double)T_ / levels < residentSweep_[0][0];
    // one instance: the backward tickets with the operator's row blocks of
    // each tile after the tiles it reads (its own and its six neighbours, at
    // most one backward level later) and a few more levels of slack, so that
    // a workgroup drawing one rarely waits
    TF_ = 0;
    if (geoms.size() == 1) {
        const TileGeom& g = geoms[0];
        const int L = g.TX + g.TY + g.TZ - 3;       // backward tile levels 0..L
        std::vector<std::vector<int>> byLevel(L + 1);
        for (int k = T_ - 1; k >= 0; --k) {         // the backward sweep's order within a level
            int TI = ord[k].tile % g.TX, TJ = (ord[k].tile / g.TX) % g.TY, TK = ord[k].tile / (g.TX * g.TY);
            byLevel[(g.TX - 1 - TI) + (g.TY - 1 - TJ) + (g.TZ - 1 - TK)].push_back(ord[k].tile);
        }
        constexpr int kSlack = 3;
        std::vector<TileTask> f;
        for (int l = 0; l <= L; ++l) {
            for (int t : byLevel[l]) f.push_back(TileTask{0, t});
            const int sl = l - 1 - kSlack;
            if (sl >= 0)
                for (int t : byLevel[sl]) f.push_back(TileTask{-1, t});
        }
        for (int sl = std::max(0, L - kSlack); sl <= L; ++sl)
            for (int t : byLevel[sl]) f.push_back(TileTask{-1, t});
        TF_ = (int)f.size();
        orderFused_.upload(f.data(), f.size(), s);
        tileDone_.alloc(T_);
        OM_HIP(hipMemsetAsync(tileDone_.p, 0xFF, T_ * sizeof(int), s));
        epochDev_.alloc(1);
        OM_HIP(hipMemsetAsync(epochDev_.p, 0, sizeof(unsigned int), s));
    }
    dcache_.clear();
    ++generation_;
}

void TiledIluSet::resetExchange(hipStream_t s)
{
    if (xch64_.p && xch64_.n)
        hipLaunchKernelGGL(k_fill_empty64, dim3((unsigned)((xch64_.n + 255) / 256)), dim3(256), 0, s,
                           (unsigned long long*)xch64_.p, (long)xch64_.n);
    if (xch32_.p && xch32_.n)
        hipLaunchKernelGGL(k_fill_empty32, dim3((unsigned)((xch32_.n + 255) / 256)), dim3(256), 0, s,
                           (unsigned int*)xch32_.p, (long)xch32_.n);
    OM_HIP(hipGetLastError());
}

template <class S>
const IluInst* TiledIluSet::upload(const std::vector<IluInst>& insts, hipStream_t s)
{
    std::vector<IluInst> v(insts);
    S* xbase = nullptr;
    if constexpr (sizeof(S) == 8) {
        xbase = (S*)xch64_.p;
    } else {
        if (!xch32_.p) {
            xch32_.alloc(std::max<long>(9 * cellOff_.back(), 1));
            hipLaunchKernelGGL(k_fill_empty32, dim3((unsigned)((xch32_.n + 255) / 256)), dim3(256), 0, s,
                               (unsigned int*)xch32_.p, (long)xch32_.n);
            OM_HIP(hipGetLastError());
        }
        xbase = (S*)xch32_.p;
    }
    for (size_t b = 0; b < v.size(); ++b) {
        v[b].done = done_.p + tileOff_[b];
        v[b].xch = xbase + 9 * cellOff_[b];
    }
    // descriptor arrays stay resident: a solver alternates between a few
    // (input, output) vector pairs, each gets its own device copy once
    for (auto& e : dcache_)
        if (e.first.size() == v.size() && std::memcmp(e.first.data(), v.data(), v.size() * sizeof(IluInst)) == 0)
            return e.second->p;
    if (dcache_.size() >= 16) {
        OM_HIP(hipStreamSynchronize(s));
        dcache_.clear();
        ++generation_;
    }
    auto buf = std::make_unique<DevBuf<IluInst>>();
    buf->upload(v.data(), v.size(), s);
    dcache_.emplace_back(v, std::move(buf));
    return dcache_.back().second->p;
}

bool TiledIluSet::takeError(hipStream_t s)
{
    int e = 0;
    OM_HIP(hipMemcpyAsync(&e, err_.p, sizeof(int), hipMemcpyDeviceToHost, s));
    OM_HIP(hipStreamSynchronize(s));
    if (!e) return false;
    // a launch that ran out of its waits leaves slots filled: empty them and the flag
    resetExchange(s);
    OM_HIP(hipMemsetAsync(err_.p, 0, sizeof(int), s));
    OM_HIP(hipStreamSynchronize(s));
    return true;
}

void TiledIluSet::raiseError(hipStream_t s)
{
    const int one = 1;
    OM_HIP(hipMemcpyAsync(err_.p, &one, sizeof(int), hipMemcpyHostToDevice, s));
    OM_HIP(hipStreamSynchronize(s));
}

void TiledIluSet::checkError(hipStream_t s)
{
    int e = 0;
    OM_HIP(hipMemcpyAsync(&e, err_.p, sizeof(int), hipMemcpyDeviceToHost, s));
    OM_HIP(hipStreamSynchronize(s));
    if (e) {
        // a launch that ran out of its waits leaves slots filled: empty them and the flag
        resetExchange(s);
        OM_HIP(hipMemsetAsync(err_.p, 0, sizeof(int), s));
        OM_HIP(hipStreamSynchronize(s));
        throw std::runtime_error("tiled ILU sweep: dependency wait exceeded its bound");
    }
}

void TiledIluSet::factor(const std::vector<IluInst>& insts, int* fail, hipStream_t s)
{
    if (T_ == 0) return;
    const IluInst* di = upload<double>(insts, s);
    long long b = 0;
    int ep = ++epoch_;
    const TileTask* od = order_.p;
    int T = T_;
    unsigned long long* ct = counter_.p;
    int* er = err_.p;
    unsigned int* ex = exits_.p;
    void* args[] = {(void*)&di, (void*)&od, &T, &ct, &b, &ep, &er, (void*)&fail, (void*)&ex};
    OM_HIP(hipLaunchKernel((const void*)k_ilu_factor_set, dim3(gridFactor_), dim3(kSweepThreads), args, 0, s));
}

template <class S>
bool TiledIluSet::apply(const std::vector<IluInst>& insts, hipStream_t s, const SpmvFuse* fuse)
{
    if (T_ == 0) return false;
    const IluInst* di = upload<S>(insts, s);
    unsigned long long* ct = counter_.p;
    int* er = err_.p;
    const int prec = sizeof(S) == 4 ? 1 : 0;
    const bool stream = streaming();
    // the operator's rows stage the tile's one-cell box of v in the sweep's LDS image
    const bool fused = fuse && stream && insts.size() == 1 && TF_ > 0 &&
                       (insts[0].g.BX + 2) * (insts[0].g.BY + 2) * (insts[0].g.BZ + 2) <= kPadMax + 1;
    unsigned int* ex = exits_.p;
    int nInst = (int)insts.size();
    for (int sweep = 0; sweep < 2; ++sweep) {
        const TileTask* od = order_.p;
        int T = T_;
        if (stream) {
            SpmvFuse fz{};
            const void* k = sweep == 0 ? (const void*)k_ilu_sweep_set<S, false, false>
                                       : (const void*)k_ilu_sweep_set<S, true, false>;
            int grid = std::min(T_, residentSweep_[prec][sweep]);
            if (sweep == 1 && fused) {
                fz = *fuse;
                fz.tileDone = tileDone_.p;
                fz.epoch = epochDev_.p;
                od = orderFused_.p;
                T = TF_;
                k = (const void*)k_ilu_sweep_set<S, true, true>;
                grid = std::min(T_, residentFused_[prec]);
            }
            void* args[] = {(void*)&di, (void*)&od, &T, &ct, &er, (void*)&ex, &nInst, &fz};
            OM_HIP(hipLaunchKernel(k, dim3(grid), dim3(kSweepThreads), args, 0, s));
        } else {
            const void* k = sweep == 0 ? (const void*)k_ilu_fwd_batch<S> : (const void*)k_ilu_bwd_batch<S>;
            const int grid = std::min(T_, residentBatch_[prec][sweep]);
            void* args[] = {(void*)&di, (void*)&od, &T, &ct, &er, (void*)&ex, &nInst};
            OM_HIP(hipLaunchKernel(k, dim3(grid), dim3(kSweepThreads), args, 0, s));
        }
    }
    return fused;
}
template bool TiledIluSet::apply<double>(const std::vector<IluInst>&, hipStream_t, const SpmvFuse*);
template bool TiledIluSet::apply<float>(const std::vector<IluInst>&, hipStream_t, const SpmvFuse*);

// ---------------------------------------------------------------------------
namespace {
std::mutex gTunedMutex;
std::map<std::array<int, 4>, TunedIlu> gTuned;
}  // namespace

bool lookupTunedIlu(int nx, int ny, int nz, int precision, TunedIlu& out)
{
    std::lock_guard<std::mutex> lock(gTunedMutex);
    auto it = gTuned.find({nx, ny, nz, precision});
    if (it == gTuned.end()) return false;
    out = it->second;
    return true;
}

void storeTunedIlu(int nx, int ny, int nz, int precision, const TunedIlu& t)
{
    std::lock_guard<std::mutex> lock(gTunedMutex);
    gTuned[{nx, ny, nz, precision}] = t;
}

// ---------------------------------------------------------------------------
// single reservoir: an ensemble of one
TiledIlu::TiledIlu(int nx, int ny, int nz, hipStream_t s) : set_(new TiledIluSet(s))
{
    g = makeTileGeom(nx, ny, nz);
    set_->setInstances({g}, s);
}

TiledIlu::TiledIlu(const TileGeom& shape, hipStream_t s, int sweepMode) : set_(new TiledIluSet(s))
{
    g = shape;
    set_->setInstances({g}, s);
    set_->setSweepMode(sweepMode);
}

void TiledIlu::setSweepMode(int mode) { set_->setSweepMode(mode); }
bool TiledIlu::streaming() const { return set_->streaming(); }

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
bool TiledIlu::apply(const int* rp, const int* diag, const S* lu, const P* dinv, const S* d, S* v, hipStream_t s,
                     const int* active, const int* ci, const S* A, S* y)
{
    static_assert(sizeof(P) == 8, "pivots are kept in FP64");
    IluInst I{};
    I.g = g; I.rp = rp; I.diag = diag; I.lu = lu; I.dinv = (const double*)dinv; I.d = d; I.v = v;
    I.active = active;
    if (!A || !y || !ci) {
        set_->apply<S>({I}, s);
        return false;
    }
    SpmvFuse fz{};
    fz.rp = rp; fz.ci = ci; fz.a = A; fz.y = y;
    return set_->apply<S>({I}, s, &fz);
}
template <class S, class P>
void TiledIlu::primeApply(const int* rp, const int* diag, const S* lu, const P* dinv, const S* d, S* v, hipStream_t s,
                          const int* active)
{
    IluInst I{};
    I.g = g; I.rp = rp; I.diag = diag; I.lu = lu; I.dinv = (const double*)dinv; I.d = d; I.v = v;
    I.active = active;
    set_->prime<S>({I}, s);
}
template void TiledIlu::primeApply<double, double>(const int*, const int*, const double*, const double*,
                                                   const double*, double*, hipStream_t, const int*);
template void TiledIlu::primeApply<float, double>(const int*, const int*, const float*, const double*, const float*,
                                                  float*, hipStream_t, const int*);

long TiledIlu::generation() const { return set_->generation(); }

template bool TiledIlu::apply<double, double>(const int*, const int*, const double*, const double*, const double*,
                                              double*, hipStream_t, const int*, const int*, const double*, double*);
template bool TiledIlu::apply<float, double>(const int*, const int*, const float*, const double*, const float*,
                                             float*, hipStream_t, const int*, const int*, const float*, float*);

}  // namespace omodels
