// ensemble.h -- per-instance device descriptors of the ensemble engine: many
// reservoir realizations of different sizes assembled and solved by the same
// launches, each kernel block finding its instance through a table.
#pragma once

#include <memory>
#include <utility>

#include "omodels_internal.h"
#include "ilu_tiled.h"

namespace omodels {

constexpr int kEnsChunk = 256;      // cells per block of the cell-parallel ensemble kernels
constexpr int kEnsDotCells = 1024;  // cells per workgroup of the ensemble's scalar products

// vectors of the Krylov solver (per instance, FP64 and FP32 sets)
enum EnsVec : int { kVb = 0, kVx, kVr, kVp, kVv, kVy, kVz, kVr64, kNumVec };

struct EnsInst {
    BoGrid G;
    int N, nnzb, nsrc, kChunk;
    int afDia;                       // 1: Af holds the structured layout [7][N][9] (two-colour ILU)
    int gx, gy, gz, nAsm;            // assembly grid of this instance
    double storefac;                 // cell volume / dt of the current solve
    const double* trans;
    const double* poro;
    double* pv0;
    uint8_t* mean0;
    uint8_t* wasSw;
    double* st1;
    double* res;
    double* jac;
    const int* rowPtr;
    const int* colIdx;
    const int* diagOff;
    const int* nbOff;
    const int* srcCell;
    const int* srcType;
    const double* srcA;
    const double* srcB;
    const int2* srcGroups;           // sources grouped by cell: {first, count} into srcOrder
    const int* srcOrder;
    double* lu64;
    double* dinv64;
    float* Af;
    float* luf;
    double* dv[kNumVec];
    float* fv[kNumVec];
    int* counters;                   // [0] switched, [1] non-finite update
};

// one tile of one instance, in ticket order
struct TileTask {
    int inst, tile;
};

// per-instance view of the tiled ILU kernels
struct IluInst {
    TileGeom g;
    const int* rp;
    const int* diag;
    const void* lu;       // S*
    void* luw;            // factorization output (double*)
    const double* dinv;
    double* dinvw;
    const void* d;        // S*
    void* v;              // S*
    int* done;            // flags of this instance's tiles (factorization)
    const int* active;    // nullptr: always active
    void* xch;            // sweep exchange slots toward the x, y, z neighbour tiles: three [N][3]
                          // arrays of S back to back (N = cells of the instance, 9 N < 2^31)
};

// The operator application that follows a preconditioner application, y = A v,
// computed inside the backward sweep of one instance: tile T's rows once the
// sweep has stored the results of T and of its -x, -y and -z neighbour tiles.
struct SpmvFuse {
    const int* rp;
    const int* ci;
    const void* a;              // S*, the operator's blocks
    void* y;                    // S*, the product
    int* tileDone;              // per tile: the launch's epoch once the tile's results are stored
    unsigned int* epoch;        // advanced by the last workgroup leaving each launch
};

// Multi-instance tiled ILU: factor / sweeps over all tiles of all instances,
// ordered by tile level so that dependencies hold smaller tickets.
class TiledIluSet {
public:
    explicit TiledIluSet(hipStream_t s);
    // (re)build the ticket order for instances with the given geometries
    void setInstances(const std::vector<TileGeom>& geoms, hipStream_t s);
    void factor(const std::vector<IluInst>& insts, int* fail, hipStream_t s);
    // fuse: also y = A v in the backward sweep when the sweeps stream and there
    // is one instance; returns whether it did
    template <class S>
    bool apply(const std::vector<IluInst>& insts, hipStream_t s, const SpmvFuse* fuse = nullptr);
    void checkError(hipStream_t s);
    // whether a dependency wait ran into its bound since the last call; the
    // flag and the exchange slots are reset either way (no throw)
    bool takeError(hipStream_t s);
    // set the error flag from the host, as an exhausted wait does: the waits of
    // the following launches give up at once (tests of the recovery)
    void raiseError(hipStream_t s);
    // make the device descriptor of these instances resident (before a stream
    // capture, which must not record the upload); generation() changes
    // whenever resident descriptors are dropped
    template <class S>
    void prime(const std::vector<IluInst>& insts, hipStream_t s) { (void)upload<S>(insts, s); }
    long generation() const { return generation_; }
    // sweep form: 0 chosen from the tiles per tile level, 1 streaming, 2 batched
    void setSweepMode(int mode) { mode_ = mode; }
    bool streaming() const { return mode_ == 1 || (mode_ == 0 && autoStream_); }
    int totalTiles() const { return T_; }
    const std::vector<int>& tileOffsets() const { return tileOff_; }

private:
    template <class S>
    const IluInst* upload(const std::vector<IluInst>& insts, hipStream_t s);
    void resetExchange(hipStream_t s);
    DevBuf<TileTask> order_;
    DevBuf<TileTask> orderFused_;        // one instance: backward tiles and the operator's row blocks
    DevBuf<int> tileDone_;
    DevBuf<unsigned int> epochDev_;
    int TF_ = 0;
    std::vector<long> cellOff_;          // first cell of each instance in the exchange arrays
    DevBuf<double> xch64_;               // [9 * cells]: three [N][3] slot arrays per instance
    DevBuf<float> xch32_;
    std::vector<std::pair<std::vector<IluInst>, std::unique_ptr<DevBuf<IluInst>>>> dcache_;
    DevBuf<int> done_, err_;
    DevBuf<unsigned long long> counter_;
    DevBuf<unsigned int> exits_;
    std::vector<int> tileOff_;
    long generation_ = 0;
    int epoch_ = 0, T_ = 0, gridFactor_ = 0, residentFactor_ = 0;
    int residentSweep_[2][2] = {};   // streaming sweeps: [FP64, FP32][forward, backward]
    int residentBatch_[2][2] = {};   // batched sweeps
    int residentFused_[2] = {};      // backward sweep with the operator application
    int mode_ = 0;
    bool autoStream_ = true;
};

}  // namespace omodels
