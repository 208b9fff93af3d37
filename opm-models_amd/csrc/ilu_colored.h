// ilu_colored.h -- block ILU(0) in a two-colour tile ordering of the 7-point
// structured grid (see ilu_colored.hip).
#pragma once

#include <vector>

#include "omodels_internal.h"

namespace omodels {

// Tiles of kColTX x kColTY x kColTZ cells (one wave, one lane per cell),
// coloured (tx + ty + tz) mod 2; colour 0 comes first, cells inside a tile in
// natural order.  Elongated in z, the direction of the strongest coupling of
// the black-oil grids (thin layers): at 500k cells the first Newton system
// needs 40 BiCGStab iterations against 30 in natural order and 65 with
// red-black cells.
constexpr int kColTX = 4, kColTY = 2, kColTZ = 8;

struct ColGeom {
    int nx, ny, nz, ntx, nty;     // grid and tiles along x and y
};
ColGeom colGeomOf(int nx, int ny, int nz);

// one reservoir of a ColoredIluSet
struct ColInst {
    ColGeom g;
    const int* rp;
    const void* A;        // the matrix (FP64 for the factorization; S for the sweeps)
    double* dinv;         // pivots
    const void* d;        // S*, right-hand side of an application
    void* v;              // S*, its result
    const int* active;    // nullptr, or a device flag: the member is skipped while it is 0
};

// The same preconditioner for many reservoirs of different sizes in the same
// launches (work items = (member, tile) per colour).
class ColoredIluSet {
public:
    void setInstances(const std::vector<ColGeom>& geoms, hipStream_t s);
    void factor(const std::vector<ColInst>& insts, int* fail, hipStream_t s);
    // slot 0 or 1: two applications with different vectors keep their own
    // descriptors (descriptor slots 1 and 2; slot 0 is the factorization's)
    template <class S>
    void apply(const std::vector<ColInst>& insts, int slot, hipStream_t s);
    // make the slot's device descriptors those of insts (an application
    // replayed from a graph uses them without the host's check)
    void prime(const std::vector<ColInst>& insts, int slot, hipStream_t s) { (void)upload(insts, 1 + slot, s); }

private:
    const ColInst* upload(const std::vector<ColInst>& insts, int slot, hipStream_t s);
    int nInst_ = 0;
    int nWork_[2] = {0, 0};
    DevBuf<int2> work_[2];
    DevBuf<ColInst> desc_[3];
    std::vector<ColInst> host_[3];
};

// dinv[N][9] = inverses of the diagonal blocks of A (slot 3 of a [7][N][9]
// structured layout, rows in any order); fail |= 2 for a singular block
void launchDiaBlockInverse(long N, const double* A, double* dinv, int* fail, hipStream_t s);

class ColoredIlu {
public:
    ColoredIlu(int nx, int ny, int nz, hipStream_t s);
    ColoredIlu(const ColoredIlu&) = delete;
    ColoredIlu& operator=(const ColoredIlu&) = delete;
    // pivots of A (FP64): dinv[N][9] = inverses of D, the diagonal of the factor
    // (D_I = A_II - sum_{K lower} A_IK D_K^-1 A_KI); fail: set to 2 for a
    // singular pivot.  rp: the sorted BCRS rows of A, or nullptr when A is in
    // the structured layout [7][N][9] (slots -z, -y, -x, I, +x, +y, +z) --
    // likewise for apply.
    void factor(const int* rp, const double* A, double* dinv, int* fail, hipStream_t s);
    // v = M^-1 d with M = (D + L_A) D^-1 (D + U_A), the off-diagonal blocks of
    // A (FP64, or its FP32 copy) and the FP64 pivots; v *= relax at the end
    // when relax != 1.  run: nullptr or a device flag, the application is
    // skipped while it is 0.  tmp: scratch of v's size.
    template <class S>
    void apply(const int* rp, const S* A, const double* dinv, const S* d, S* v, double relax, hipStream_t s,
               const int* run = nullptr);
    int nx() const { return nx_; }
    int ny() const { return ny_; }
    int nz() const { return nz_; }

private:
    int nx_, ny_, nz_;
    int nTiles_[2] = {0, 0};
    DevBuf<int> tiles_[2];     // tile indices of each colour
};

}  // namespace omodels
