// amg.h -- aggregation AMG of a 3x3 system on the 7-point grid: the
// preconditioner of precond = 2, in place of the reference's
// ParallelAmgBackend (parallelamgbackend.hh:85-316; see amg.hip).
#pragma once

#include <array>
#include <memory>
#include <vector>

#include "omodels_internal.h"

namespace omodels {

// the coarsest level is solved by a dense inverse of at most this many cells
constexpr int kAmgMaxCoarse = 32;

struct AmgParams {
    int coarsenTarget = 32;     // levels until one has at most this many cells (1..kAmgMaxCoarse)
    int pre = 2, post = 2;      // forward / backward Gauss-Seidel sweeps around the coarse correction
    double damp = 1.6;          // x += damp * P xc
    int agg[3] = {3, 3, 3};     // aggregate box in cells
};
// throws std::invalid_argument for values outside their ranges
void checkAmgParams(const AmgParams& p);
// level dimensions of an nx x ny x nz grid
std::vector<std::array<int, 3>> amgLevelDims(int nx, int ny, int nz, const AmgParams& p);

class AmgHierarchy {
public:
    // fp32: also keep FP32 copies of the level matrices for applications in
    // FP32 (the inner solve of the precision form; pivots stay FP64)
    AmgHierarchy(int nx, int ny, int nz, const AmgParams& p, bool fp32, hipStream_t s);
    AmgHierarchy(const AmgHierarchy&) = delete;
    AmgHierarchy& operator=(const AmgHierarchy&) = delete;
    // the levels from the fine matrix A0 (structured layout [7][N][9], FP64):
    // its red-black copy, Galerkin coarse operators, inverted diagonal
    // blocks, the coarsest level's dense inverse; fail |= 2 for a singular
    // diagonal block or coarsest matrix
    void setup(const double* A0, int* fail, hipStream_t s);
    // v = one V-cycle applied to d (device vectors of 3N, FP64, or FP32 with
    // the FP32 copies); skipped while *run == 0
    template <class S>
    void apply(const S* d, S* v, const int* run, hipStream_t s);
    int levels() const { return (int)lv_.size(); }
    std::array<int, 3> dims(int l) const { return {lv_[l]->nx, lv_[l]->ny, lv_[l]->nz}; }
    // level l's matrix in the natural structured layout [7][N_l][9] (host)
    void downloadLevel(int l, double* host, hipStream_t s) const;

private:
    struct Level {
        int nx, ny, nz;
        long N, N0;                             // cells, even cells
        int ax = 1, ay = 1, az = 1;             // aggregation to the next level
        DevBuf<double> A;                       // red-black layout: even cells' rows, then odd ones'
        DevBuf<double> dinv, b, x, r;
        DevBuf<float> Af, bf, xf, rf;           // FP32 copies (fp32)
        DevBuf<int> rbOf, rbInv;                // natural cell -> red-black row, and back
        std::vector<int> hostRbOf;
        template <class S> const S* mat() const
        {
            if constexpr (sizeof(S) == 8) return A.p; else return Af.p;
        }
        template <class S> S* res()
        {
            if constexpr (sizeof(S) == 8) return r.p; else return rf.p;
        }
        template <class S> S* rhs()
        {
            if constexpr (sizeof(S) == 8) return b.p; else return bf.p;
        }
        template <class S> S* sol()
        {
            if constexpr (sizeof(S) == 8) return x.p; else return xf.p;
        }
    };
    template <class S> void cycle(size_t l, const S* b, S* x, const int* run, hipStream_t s);
    std::vector<std::unique_ptr<Level>> lv_;
    DevBuf<double> dense_;
    AmgParams p_;
    bool fp32_ = false;
    bool ready_ = false;
};

}  // namespace omodels
