// ilu_colored.hip -- block ILU(0) of the 7-point black-oil pattern in a
// two-colour tile ordering, for the forms whose bar is an accuracy against the
// FP64 result rather than the reference's natural-order bits (the FP32-inner
// solve and the ensembles).
//
// The reference's preconditioner is dune's SeqILU<.,.,.,0> in the natural
// cell order (PreconditionerWrapperILU, istlpreconditionerwrappers.hh:151-190);
// its triangular sweeps are one dependency chain of nx+ny+nz-2 cell levels,
// which bounds them by latency on a GPU.  Here the grid is cut into tiles of
// kColTX x kColTY x kColTZ cells, coloured like a chess board; all tiles of
// colour 0 come first, then all of colour 1, and the cells of a tile in
// natural order.  ILU(0) of the matrix in that order has the same form as the
// reference's: on a 7-point stencil two neighbouring cells share no third
// neighbour, so elimination only changes the diagonal blocks,
//     D_I = A_II - sum_{K lower of I} A_IK D_K^-1 A_KI,
// and M = (D + L_A) D^-1 (D + U_A) with the off-diagonal blocks of A.  A
// sweep is two launches (one per colour); a tile is one wave, one lane per
// cell, its 12 cell levels exchanged between lanes by shuffles.  Every tile of
// a colour runs at once, so a sweep is bound by the bytes it reads.
//
// Operation order (the CPU restatement follows it): per cell, the
// contributions of neighbours in other tiles in column order, then those of
// its own tile in column order; products as row sums (s = a0 x0 + a1 x1 +
// a2 x2), subtracted from the accumulator.
#include "ilu_colored.h"

#include <cstring>
#include <stdexcept>

namespace omodels {

namespace {

constexpr int kWavesPerBlock = 4;

// cell of this lane in tile t; false when outside the grid (ragged edge tiles)
__device__ __forceinline__ bool laneCell(const ColGeom& g, int t, int lane, int& gx, int& gy, int& gz, int& lx,
                                         int& ly, int& lz)
{
    const int tx = t % g.ntx, ty = (t / g.ntx) % g.nty, tz = t / (g.ntx * g.nty);
    lx = lane & 3; ly = (lane >> 2) & 1; lz = lane >> 3;
    gx = tx * kColTX + lx; gy = ty * kColTY + ly; gz = tz * kColTZ + lz;
    return gx < g.nx && gy < g.ny && gz < g.nz;
}

// Block positions of a row's neighbours in the sorted BCRS row
// (-z, -y, -x, I, +x, +y, +z where present); -1 where absent.
struct RowPos {
    long blk[6];      // by direction 0 -x, 1 +x, 2 -y, 3 +y, 4 -z, 5 +z
    long diag;
};
__device__ __forceinline__ RowPos rowPositions(const ColGeom& g, long k0, int gx, int gy, int gz)
{
    RowPos p;
    long k = k0;
    p.blk[4] = gz > 0 ? k++ : -1;
    p.blk[2] = gy > 0 ? k++ : -1;
    p.blk[0] = gx > 0 ? k++ : -1;
    p.diag = k++;
    p.blk[1] = gx + 1 < g.nx ? k++ : -1;
    p.blk[3] = gy + 1 < g.ny ? k++ : -1;
    p.blk[5] = gz + 1 < g.nz ? k++ : -1;
    return p;
}

// The same positions in the structured layout [7][N][9] (slots -z, -y, -x,
// I, +x, +y, +z; blocks of absent neighbours are zero and never read): the
// blocks of one slot of consecutive cells are consecutive in memory.
__device__ __forceinline__ RowPos diaPositions(const ColGeom& g, long I, int gx, int gy, int gz)
{
    const long N = (long)g.nx * g.ny * g.nz;
    RowPos p;
    p.blk[4] = gz > 0 ? I : -1;
    p.blk[2] = gy > 0 ? N + I : -1;
    p.blk[0] = gx > 0 ? 2 * N + I : -1;
    p.diag = 3 * N + I;
    p.blk[1] = gx + 1 < g.nx ? 4 * N + I : -1;
    p.blk[3] = gy + 1 < g.ny ? 5 * N + I : -1;
    p.blk[5] = gz + 1 < g.nz ? 6 * N + I : -1;
    return p;
}

// rp != nullptr: the sorted BCRS rows; rp == nullptr: the structured layout
__device__ __forceinline__ RowPos positionsOf(const ColGeom& g, const int* __restrict__ rp, long I, int gx, int gy,
                                             int gz)
{
    return rp ? rowPositions(g, rp[I], gx, gy, gz) : diaPositions(g, I, gx, gy, gz);
}

// whether the neighbour in direction dir lies in another tile
__device__ __forceinline__ bool crossesTile(int dir, int lx, int ly, int lz)
{
    switch (dir) {
    case 0: return lx == 0;
    case 1: return lx == kColTX - 1;
    case 2: return ly == 0;
    case 3: return ly == kColTY - 1;
    case 4: return lz == 0;
    default: return lz == kColTZ - 1;
    }
}

__device__ __forceinline__ long neighbourCell(const ColGeom& g, long I, int dir)
{
    const long nxy = (long)g.nx * g.ny;
    switch (dir) {
    case 0: return I - 1;
    case 1: return I + 1;
    case 2: return I - g.nx;
    case 3: return I + g.nx;
    case 4: return I - nxy;
    default: return I + nxy;
    }
}

// directions in column order: -z, -y, -x, +x, +y, +z
__device__ constexpr int kColOrder[6] = {4, 2, 0, 1, 3, 5};

template <class S>
__device__ __forceinline__ void subBlockVec(const S* __restrict__ blk, S x0, S x1, S x2, S (&acc)[3])
{
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        S s = blk[3 * e] * x0;
        s += blk[3 * e + 1] * x1;
        s += blk[3 * e + 2] * x2;
        acc[e] += S(-1) * s;
    }
}

__device__ __forceinline__ void matmul3(const double* A, const double* B, double* C)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double s = A[3 * i] * B[j];
            s += A[3 * i + 1] * B[3 + j];
            s += A[3 * i + 2] * B[6 + j];
            C[3 * i + j] = s;
        }
}

__device__ __forceinline__ bool invert3(double* a)
{
    const double c00 = a[4] * a[8] - a[5] * a[7];
    const double c01 = a[5] * a[6] - a[3] * a[8];
    const double c02 = a[3] * a[7] - a[4] * a[6];
    const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (!(det != 0.0) || !isfinite(det)) return false;
    const double r = 1.0 / det;
    const double c10 = a[2] * a[7] - a[1] * a[8];
    const double c11 = a[0] * a[8] - a[2] * a[6];
    const double c12 = a[1] * a[6] - a[0] * a[7];
    const double c20 = a[1] * a[5] - a[2] * a[4];
    const double c21 = a[2] * a[3] - a[0] * a[5];
    const double c22 = a[0] * a[4] - a[1] * a[3];
    a[0] = c00 * r; a[1] = c10 * r; a[2] = c20 * r;
    a[3] = c01 * r; a[4] = c11 * r; a[5] = c21 * r;
    a[6] = c02 * r; a[7] = c12 * r; a[8] = c22 * r;
    return true;
}

// in-tile lower neighbours by lane offset: -z (8), -y (4), -x (1)
constexpr int kInLowerDir[3] = {4, 2, 0};
constexpr int kInLowerLane[3] = {8, 4, 1};
// in-tile upper neighbours in column order: +x (1), +y (4), +z (8)
constexpr int kInUpperDir[3] = {1, 3, 5};
constexpr int kInUpperLane[3] = {1, 4, 8};

__device__ __forceinline__ bool inTileLower(int q, int lx, int ly, int lz)
{
    return q == 0 ? lz > 0 : (q == 1 ? ly > 0 : lx > 0);
}
__device__ __forceinline__ bool inTileUpper(int q, int lx, int ly, int lz)
{
    return q == 0 ? lx < kColTX - 1 : (q == 1 ? ly < kColTY - 1 : lz < kColTZ - 1);
}

constexpr int kLevels = (kColTX - 1) + (kColTY - 1) + (kColTZ - 1) + 1;

// Factorization of the tiles of one colour: pivots D_I^-1 (FP64).
// (one wave: every lane of the wave calls it for the same tile)
__device__ __forceinline__ void colFactorTile(const ColGeom& g, int color, int tile, const int* __restrict__ rp,
                                              const double* __restrict__ A, double* __restrict__ dinv,
                                              int* __restrict__ fail)
{
    const int lane = (int)(threadIdx.x & 63u);
    int gx, gy, gz, lx, ly, lz;
    const bool valid = laneCell(g, tile, lane, gx, gy, gz, lx, ly, lz);
    const long I = ((long)gz * g.ny + gy) * g.nx + gx;
    double D[9];
    RowPos pos{};
    if (valid) {
        pos = positionsOf(g, rp, I, gx, gy, gz);
#pragma unroll
        for (int q = 0; q < 9; ++q) D[q] = A[9 * pos.diag + q];
        // neighbours in tiles of the earlier colour (colour 1 only), column order
        if (color == 1) {
#pragma unroll
            for (int o = 0; o < 6; ++o) {
                const int dir = kColOrder[o];
                if (pos.blk[dir] < 0 || !crossesTile(dir, lx, ly, lz)) continue;
                const long K = neighbourCell(g, I, dir);
                // A_KI: the block of row K in column I, at K's opposite-direction position
                int kx = (int)(K % g.nx), ky = (int)((K / g.nx) % g.ny), kz = (int)(K / ((long)g.nx * g.ny));
                const RowPos pk = positionsOf(g, rp, K, kx, ky, kz);
                const long pKI = pk.blk[dir ^ 1];
                double P[9], T[9];
                matmul3(A + 9 * pos.blk[dir], dinv + 9 * K, P);
                matmul3(P, A + 9 * pKI, T);
#pragma unroll
                for (int q = 0; q < 9; ++q) D[q] -= T[q];
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < 9; ++q) D[q] = 0.0;
    }
    // in-tile lower neighbours: their A_IK and A_KI blocks
    double Dinv[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) Dinv[q] = 0.0;
    const int lev = lx + ly + lz;
    for (int l = 0; l < kLevels; ++l) {
        // the pivots of the three possible lower neighbours of this tile
        double nb[3][9];
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int c = 0; c < 9; ++c) nb[q][c] = __shfl(Dinv[c], lane - kInLowerLane[q] < 0 ? lane : lane - kInLowerLane[q], 64);
        if (valid && lev == l) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                if (!inTileLower(q, lx, ly, lz)) continue;
                const int dir = kInLowerDir[q];
                const long K = neighbourCell(g, I, dir);
                int kx = (int)(K % g.nx), ky = (int)((K / g.nx) % g.ny), kz = (int)(K / ((long)g.nx * g.ny));
                const RowPos pk = positionsOf(g, rp, K, kx, ky, kz);
                double P[9], T[9];
                matmul3(A + 9 * pos.blk[dir], nb[q], P);
                matmul3(P, A + 9 * pk.blk[dir ^ 1], T);
#pragma unroll
                for (int c = 0; c < 9; ++c) D[c] -= T[c];
            }
#pragma unroll
            for (int c = 0; c < 9; ++c) Dinv[c] = D[c];
            if (!invert3(Dinv)) atomicOr(fail, 2);
        }
    }
    if (valid) {
#pragma unroll
        for (int c = 0; c < 9; ++c) dinv[9 * I + c] = Dinv[c];
    }
}

// Forward sweep of one colour: w_I = D_I^-1 (d_I - sum_{K lower} A_IK w_K), w into v.
template <class S>
__device__ __forceinline__ void colForwardTile(const ColGeom& g, int color, int tile, const int* __restrict__ rp,
                                               const S* __restrict__ A, const double* __restrict__ dinv,
                                               const S* __restrict__ d, S* __restrict__ v)
{
    const int lane = (int)(threadIdx.x & 63u);
    int gx, gy, gz, lx, ly, lz;
    const bool valid = laneCell(g, tile, lane, gx, gy, gz, lx, ly, lz);
    const long I = ((long)gz * g.ny + gy) * g.nx + gx;
    S acc[3] = {S(0), S(0), S(0)};
    S blk[3][9];
    double Di[9];
    if (valid) {
        const RowPos pos = positionsOf(g, rp, I, gx, gy, gz);
#pragma unroll
        for (int e = 0; e < 3; ++e) acc[e] = d[3 * I + e];
        if (color == 1) {
#pragma unroll
            for (int o = 0; o < 6; ++o) {
                const int dir = kColOrder[o];
                if (pos.blk[dir] < 0 || !crossesTile(dir, lx, ly, lz)) continue;
                const long K = neighbourCell(g, I, dir);
                subBlockVec(A + 9 * pos.blk[dir], v[3 * K], v[3 * K + 1], v[3 * K + 2], acc);
            }
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const bool has = inTileLower(q, lx, ly, lz);
            const long at = has ? pos.blk[kInLowerDir[q]] : pos.diag;   // a valid address either way
#pragma unroll
            for (int c = 0; c < 9; ++c) blk[q][c] = has ? A[9 * at + c] : S(0);
        }
#pragma unroll
        for (int c = 0; c < 9; ++c) Di[c] = dinv[9 * I + c];
    }
    S wI[3] = {S(0), S(0), S(0)};
    const int lev = lx + ly + lz;
    for (int l = 0; l < kLevels; ++l) {
        S nb[3][3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int src = lane - kInLowerLane[q] < 0 ? lane : lane - kInLowerLane[q];
#pragma unroll
            for (int e = 0; e < 3; ++e) nb[q][e] = __shfl(wI[e], src, 64);
        }
        if (valid && lev == l) {
#pragma unroll
            for (int q = 0; q < 3; ++q)
                if (inTileLower(q, lx, ly, lz)) subBlockVec(blk[q], nb[q][0], nb[q][1], nb[q][2], acc);
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                double o = Di[3 * e] * (double)acc[0];
                o += Di[3 * e + 1] * (double)acc[1];
                o += Di[3 * e + 2] * (double)acc[2];
                wI[e] = (S)(0.0 + o);
            }
        }
    }
    if (valid) {
#pragma unroll
        for (int e = 0; e < 3; ++e) v[3 * I + e] = wI[e];
    }
}

// Backward sweep of one colour, in place: x_I = w_I - D_I^-1 sum_{J upper} A_IJ x_J.
template <class S>
__device__ __forceinline__ void colBackwardTile(const ColGeom& g, int color, int tile, const int* __restrict__ rp,
                                                const S* __restrict__ A, const double* __restrict__ dinv,
                                                S* __restrict__ v)
{
    const int lane = (int)(threadIdx.x & 63u);
    int gx, gy, gz, lx, ly, lz;
    const bool valid = laneCell(g, tile, lane, gx, gy, gz, lx, ly, lz);
    const long I = ((long)gz * g.ny + gy) * g.nx + gx;
    S acc[3] = {S(0), S(0), S(0)};
    S wOwn[3] = {S(0), S(0), S(0)};
    S blk[3][9];
    double Di[9];
    bool hasUp[3] = {false, false, false};
    if (valid) {
        const RowPos pos = positionsOf(g, rp, I, gx, gy, gz);
#pragma unroll
        for (int e = 0; e < 3; ++e) wOwn[e] = v[3 * I + e];
        // neighbours in tiles of the later colour (colour 0 only), column order
        if (color == 0) {
#pragma unroll
            for (int o = 0; o < 6; ++o) {
                const int dir = kColOrder[o];
                if (pos.blk[dir] < 0 || !crossesTile(dir, lx, ly, lz)) continue;
                const long J = neighbourCell(g, I, dir);
                subBlockVec(A + 9 * pos.blk[dir], v[3 * J], v[3 * J + 1], v[3 * J + 2], acc);
            }
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            hasUp[q] = inTileUpper(q, lx, ly, lz) && pos.blk[kInUpperDir[q]] >= 0;
            const long at = hasUp[q] ? pos.blk[kInUpperDir[q]] : pos.diag;
#pragma unroll
            for (int c = 0; c < 9; ++c) blk[q][c] = hasUp[q] ? A[9 * at + c] : S(0);
        }
#pragma unroll
        for (int c = 0; c < 9; ++c) Di[c] = dinv[9 * I + c];
    }
    S xI[3] = {S(0), S(0), S(0)};
    const int lev = lx + ly + lz;
    for (int l = kLevels - 1; l >= 0; --l) {
        S nb[3][3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int src = lane + kInUpperLane[q] > 63 ? lane : lane + kInUpperLane[q];
#pragma unroll
            for (int e = 0; e < 3; ++e) nb[q][e] = __shfl(xI[e], src, 64);
        }
        if (valid && lev == l) {
#pragma unroll
            for (int q = 0; q < 3; ++q)
                if (hasUp[q]) subBlockVec(blk[q], nb[q][0], nb[q][1], nb[q][2], acc);
            // acc = -sum A_IJ x_J; x = w + D^-1 acc
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                double o = Di[3 * e] * (double)acc[0];
                o += Di[3 * e + 1] * (double)acc[1];
                o += Di[3 * e + 2] * (double)acc[2];
                xI[e] = (S)((double)wOwn[e] + (0.0 + o));
            }
        }
    }
    if (valid) {
#pragma unroll
        for (int e = 0; e < 3; ++e) v[3 * I + e] = xI[e];
    }
}

// inverted diagonal blocks of a structured-layout matrix; fail |= 2 for a
// singular block
__global__ void k_dia_block_inverse(long N, const double* __restrict__ A, double* __restrict__ dinv,
                                    int* __restrict__ fail)
{
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    double D[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) D[c] = A[9 * (3 * N + I) + c];
    if (!invert3(D)) atomicOr(fail, 2);
#pragma unroll
    for (int c = 0; c < 9; ++c) dinv[9 * I + c] = D[c];
}

// one reservoir: the tiles of one colour, one wave each
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_col_factor(ColGeom g, int color,
    const int* __restrict__ tiles, int nTiles, const int* __restrict__ rp, const double* __restrict__ A,
    double* __restrict__ dinv, int* __restrict__ fail)
{
    const int w = blockIdx.x * kWavesPerBlock + (int)(threadIdx.x >> 6);
    if (w >= nTiles) return;                     // wave-uniform
    colFactorTile(g, color, tiles[w], rp, A, dinv, fail);
}
template <class S>
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_col_forward(ColGeom g, int color,
    const int* __restrict__ tiles, int nTiles, const int* __restrict__ rp, const S* __restrict__ A,
    const double* __restrict__ dinv, const S* __restrict__ d, S* __restrict__ v, const int* __restrict__ run)
{
    if (run && !*run) return;
    const int w = blockIdx.x * kWavesPerBlock + (int)(threadIdx.x >> 6);
    if (w >= nTiles) return;
    colForwardTile<S>(g, color, tiles[w], rp, A, dinv, d, v);
}
template <class S>
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_col_backward(ColGeom g, int color,
    const int* __restrict__ tiles, int nTiles, const int* __restrict__ rp, const S* __restrict__ A,
    const double* __restrict__ dinv, S* __restrict__ v, const int* __restrict__ run)
{
    if (run && !*run) return;
    const int w = blockIdx.x * kWavesPerBlock + (int)(threadIdx.x >> 6);
    if (w >= nTiles) return;
    colBackwardTile<S>(g, color, tiles[w], rp, A, dinv, v);
}

// several reservoirs: work items (member, tile) of one colour, one wave each;
// members whose flag is 0 are skipped
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_col_factor_set(const ColInst* __restrict__ insts,
    const int2* __restrict__ work, int nWork, int color, int* __restrict__ fail)
{
    const int w = blockIdx.x * kWavesPerBlock + (int)(threadIdx.x >> 6);
    if (w >= nWork) return;
    const int2 it = work[w];
    const ColInst& I = insts[it.x];
    if (I.active && !*I.active) return;
    colFactorTile(I.g, color, it.y, I.rp, (const double*)I.A, I.dinv, fail);
}
template <class S>
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_col_forward_set(const ColInst* __restrict__ insts,
    const int2* __restrict__ work, int nWork, int color)
{
    const int w = blockIdx.x * kWavesPerBlock + (int)(threadIdx.x >> 6);
    if (w >= nWork) return;
    const int2 it = work[w];
    const ColInst& I = insts[it.x];
    if (I.active && !*I.active) return;
    colForwardTile<S>(I.g, color, it.y, I.rp, (const S*)I.A, I.dinv, (const S*)I.d, (S*)I.v);
}
template <class S>
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_col_backward_set(const ColInst* __restrict__ insts,
    const int2* __restrict__ work, int nWork, int color)
{
    const int w = blockIdx.x * kWavesPerBlock + (int)(threadIdx.x >> 6);
    if (w >= nWork) return;
    const int2 it = work[w];
    const ColInst& I = insts[it.x];
    if (I.active && !*I.active) return;
    colBackwardTile<S>(I.g, color, it.y, I.rp, (const S*)I.A, I.dinv, (S*)I.v);
}

template <class S>
__global__ void k_col_scale(long m, S* __restrict__ v, double w, const int* __restrict__ run)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || (run && !*run)) return;
    v[i] = (S)(w * (double)v[i]);
}

}  // namespace

ColGeom colGeomOf(int nx, int ny, int nz)
{
    ColGeom g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.ntx = (nx + kColTX - 1) / kColTX;
    g.nty = (ny + kColTY - 1) / kColTY;
    return g;
}

// tiles of each colour of a grid
static void colourLists(const ColGeom& g, std::vector<int> (&lists)[2])
{
    const int ntz = (g.nz + kColTZ - 1) / kColTZ;
    for (int tz = 0; tz < ntz; ++tz)
        for (int ty = 0; ty < g.nty; ++ty)
            for (int tx = 0; tx < g.ntx; ++tx)
                lists[(tx + ty + tz) & 1].push_back(tx + g.ntx * (ty + g.nty * tz));
}

ColoredIlu::ColoredIlu(int nx, int ny, int nz, hipStream_t s) : nx_(nx), ny_(ny), nz_(nz)
{
    if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("colored ILU: empty grid");
    std::vector<int> lists[2];
    colourLists(colGeomOf(nx, ny, nz), lists);
    for (int c = 0; c < 2; ++c) {
        nTiles_[c] = (int)lists[c].size();
        if (nTiles_[c]) tiles_[c].upload(lists[c].data(), lists[c].size(), s);
    }
}

void ColoredIlu::factor(const int* rp, const double* A, double* dinv, int* fail, hipStream_t s)
{
    const ColGeom g = colGeomOf(nx_, ny_, nz_);
    for (int c = 0; c < 2; ++c) {
        if (!nTiles_[c]) continue;
        const int blocks = (nTiles_[c] + kWavesPerBlock - 1) / kWavesPerBlock;
        hipLaunchKernelGGL(k_col_factor, dim3(blocks), dim3(64 * kWavesPerBlock), 0, s, g, c, tiles_[c].p,
                           nTiles_[c], rp, A, dinv, fail);
    }
    OM_HIP(hipGetLastError());
}

template <class S>
void ColoredIlu::apply(const int* rp, const S* A, const double* dinv, const S* d, S* v, double relax,
                       hipStream_t s, const int* run)
{
    const ColGeom g = colGeomOf(nx_, ny_, nz_);
    for (int c = 0; c < 2; ++c) {
        if (!nTiles_[c]) continue;
        const int blocks = (nTiles_[c] + kWavesPerBlock - 1) / kWavesPerBlock;
        hipLaunchKernelGGL((k_col_forward<S>), dim3(blocks), dim3(64 * kWavesPerBlock), 0, s, g, c, tiles_[c].p,
                           nTiles_[c], rp, A, dinv, d, v, run);
    }
    for (int c = 1; c >= 0; --c) {
        if (!nTiles_[c]) continue;
        const int blocks = (nTiles_[c] + kWavesPerBlock - 1) / kWavesPerBlock;
        hipLaunchKernelGGL((k_col_backward<S>), dim3(blocks), dim3(64 * kWavesPerBlock), 0, s, g, c, tiles_[c].p,
                           nTiles_[c], rp, A, dinv, v, run);
    }
    if (relax != 1.0) {
        const long m = 3L * nx_ * ny_ * nz_;
        hipLaunchKernelGGL((k_col_scale<S>), dim3((m + 255) / 256), dim3(256), 0, s, m, v, relax, run);
    }
    OM_HIP(hipGetLastError());
}

void launchDiaBlockInverse(long N, const double* A, double* dinv, int* fail, hipStream_t s)
{
    if (N <= 0) return;
    hipLaunchKernelGGL(k_dia_block_inverse, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, N, A, dinv, fail);
    OM_HIP(hipGetLastError());
}

template void ColoredIlu::apply<double>(const int*, const double*, const double*, const double*, double*, double,
                                        hipStream_t, const int*);
template void ColoredIlu::apply<float>(const int*, const float*, const double*, const float*, float*, double,
                                       hipStream_t, const int*);

}  // namespace omodels

namespace omodels {

// ---------------------------------------------------------------------------
void ColoredIluSet::setInstances(const std::vector<ColGeom>& geoms, hipStream_t s)
{
    std::vector<int2> work[2];
    for (size_t b = 0; b < geoms.size(); ++b) {
        std::vector<int> lists[2];
        colourLists(geoms[b], lists);
        for (int c = 0; c < 2; ++c)
            for (int t : lists[c]) work[c].push_back(make_int2((int)b, t));
    }
    for (int c = 0; c < 2; ++c) {
        nWork_[c] = (int)work[c].size();
        if (nWork_[c]) work_[c].upload(work[c].data(), work[c].size(), s);
    }
    nInst_ = (int)geoms.size();
}

const ColInst* ColoredIluSet::upload(const std::vector<ColInst>& insts, int slot, hipStream_t s)
{
    if ((int)insts.size() != nInst_) throw std::logic_error("colored ILU set: instance count changed");
    DevBuf<ColInst>& d = desc_[slot];
    if (host_[slot].size() != insts.size() ||
        std::memcmp(host_[slot].data(), insts.data(), insts.size() * sizeof(ColInst)) != 0) {
        host_[slot] = insts;
        d.upload(insts.data(), insts.size(), s);
    }
    return d.p;
}

void ColoredIluSet::factor(const std::vector<ColInst>& insts, int* fail, hipStream_t s)
{
    const ColInst* di = upload(insts, 0, s);
    for (int c = 0; c < 2; ++c) {
        if (!nWork_[c]) continue;
        const int blocks = (nWork_[c] + kWavesPerBlock - 1) / kWavesPerBlock;
        hipLaunchKernelGGL(k_col_factor_set, dim3(blocks), dim3(64 * kWavesPerBlock), 0, s, di, work_[c].p,
                           nWork_[c], c, fail);
    }
    OM_HIP(hipGetLastError());
}

template <class S>
void ColoredIluSet::apply(const std::vector<ColInst>& insts, int slot, hipStream_t s)
{
    const ColInst* di = upload(insts, 1 + slot, s);
    for (int c = 0; c < 2; ++c) {
        if (!nWork_[c]) continue;
        const int blocks = (nWork_[c] + kWavesPerBlock - 1) / kWavesPerBlock;
        hipLaunchKernelGGL((k_col_forward_set<S>), dim3(blocks), dim3(64 * kWavesPerBlock), 0, s, di, work_[c].p,
                           nWork_[c], c);
    }
    for (int c = 1; c >= 0; --c) {
        if (!nWork_[c]) continue;
        const int blocks = (nWork_[c] + kWavesPerBlock - 1) / kWavesPerBlock;
        hipLaunchKernelGGL((k_col_backward_set<S>), dim3(blocks), dim3(64 * kWavesPerBlock), 0, s, di, work_[c].p,
                           nWork_[c], c);
    }
    OM_HIP(hipGetLastError());
}
template void ColoredIluSet::apply<double>(const std::vector<ColInst>&, int, hipStream_t);
template void ColoredIluSet::apply<float>(const std::vector<ColInst>&, int, hipStream_t);

}  // namespace omodels
