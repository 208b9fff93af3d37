// omodels_internal.h -- shared declarations of the MI355X host runtime.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "bo_physics.h"
#include "exact_sum.h"

namespace omodels {

// Thrown for the reference's NumericalProblem (recoverable: the time loop
// halves the step, fvbaseproblem.hh:512-557).
struct NumericalProblem : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hipCheck(hipError_t e, const char* what);
#define OM_HIP(x) ::omodels::hipCheck((x), #x)

struct BoGrid {
    int nx, ny, nz, pad;
    double dx, dy, dz, g, V;
    double area[3];       // face areas of x, y and z faces
};

// The linearizer's per-neighbour inputs given by the caller instead of the
// structured grid's own (tpfalinearizer.hh:443-452): each cell's centre depth
// (dofCenterDepth) and total volume (dofTotalVolume), the threshold pressure
// per face ([3][N] at the transmissibility's index, zeros where none), the
// time step of the linearization (storefac = volume / dt, :776-778) and the
// rock's compressibility and multipliers (BoRock).  depth == nullptr: the
// grid's own geometry and no rock terms (the kernels' plain form).
struct BoGeo {
    const double* depth;
    const double* vol;
    const double* thp;
    double dt;
    BoRock rock;
    const double* dirMob;     // [3 axes][3 phases][N] directional mobility multipliers, or nullptr
    const uint8_t* domain;    // a sub-domain's cells (nonzero), or nullptr: every cell
};

// Boundary faces with a condition, sorted by cell and face, grouped by cell
// ({first, count} per cell that has any)
struct BoBc {
    int nGroups;
    const int2* groups;
    const int* cell;
    const int* face;
    const int* type;            // 0 none, 1 rate, 2 / 3 free (Dirichlet)
    const double* rate;         // [n][3] mass rate per area, equation order
    const double* state;        // [n][3] the boundary's primary variables
    const uint8_t* meaning;     // [n]
    const double* trans;        // [n]
    const double* zFace;        // [n]
};

struct NewtonUpdateParams {
    double dpMaxRel, dsMax, oscThresh;
};

// Device buffer owning a hipMalloc'ed array.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr; n = 0;
    }
    void alloc(size_t count)
    {
        if (count == n && p) return;
        release();
        if (count) OM_HIP(hipMalloc((void**)&p, count * sizeof(T)));
        n = count;
    }
    void upload(const T* h, size_t count, hipStream_t s)
    {
        alloc(count);
        if (count) OM_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T* h, size_t count, hipStream_t s) const
    {
        if (count) OM_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

// Optional HIP-event timing of kernel classes on the stream they run on.  With
// every = k only each k-th launch of a class is bracketed by events (an event
// record costs the stream a few microseconds); launched counts all of them.
enum KernelClass : int { kClsAssemble = 0, kClsSpmv, kClsIluApply, kClsIluFactor, kClsOther, kNumClasses };
struct KernelTimer {
    bool enabled = false;
    int every = 1;
    struct Pair { hipEvent_t a, b; int cls; long idx; };
    std::vector<Pair> pending;
    std::vector<hipEvent_t> pool;
    double ms[kNumClasses] = {0, 0, 0, 0, 0};        // summed over the timed launches
    long count[kNumClasses] = {0, 0, 0, 0, 0};       // timed launches
    long launched[kNumClasses] = {0, 0, 0, 0, 0};    // all launches
    hipEvent_t cur = nullptr;
    int curCls = 0;
    long curIdx = 0;
    hipEvent_t get()
    {
        if (pool.empty()) { hipEvent_t e; OM_HIP(hipEventCreate(&e)); return e; }
        hipEvent_t e = pool.back(); pool.pop_back(); return e;
    }
    void begin(int cls, hipStream_t s)
    {
        if (!enabled) return;
        curIdx = launched[cls]++;
        if (curIdx % every != 0) { cur = nullptr; return; }
        cur = get(); curCls = cls;
        OM_HIP(hipEventRecord(cur, s));
    }
    void end(hipStream_t s)
    {
        if (!enabled || !cur) return;
        hipEvent_t b = get();
        OM_HIP(hipEventRecord(b, s));
        pending.push_back({cur, b, curCls, curIdx});
        cur = nullptr;
        if (pending.size() > 8192) collect();
    }
    void collect()
    {
        for (auto& p : pending) {
            OM_HIP(hipEventSynchronize(p.b));
            float t = 0.f;
            OM_HIP(hipEventElapsedTime(&t, p.a, p.b));
            ms[p.cls] += t; count[p.cls] += 1;
            pool.push_back(p.a); pool.push_back(p.b);
        }
        pending.clear();
    }
    void reset() { collect(); for (int i = 0; i < kNumClasses; ++i) { ms[i] = 0; count[i] = 0; launched[i] = 0; } }
    // forget the last n launches of a class (queued work that the device skipped)
    void dropLast(int cls, long n)
    {
        if (!enabled || n <= 0) return;
        const long from = launched[cls] - n;
        std::vector<Pair> keep;
        for (auto& p : pending) {
            if (p.cls == cls && p.idx >= from) { pool.push_back(p.a); pool.push_back(p.b); }
            else keep.push_back(p);
        }
        pending.swap(keep);
        launched[cls] = from;
    }
    // total time of a class over all its launches (the timed ones' average times the launch count)
    double estimatedMs(int cls) const { return count[cls] ? ms[cls] * (double)launched[cls] / count[cls] : 0.0; }
    ~KernelTimer()
    {
        for (auto& p : pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// ---- black-oil kernels (bo_kernels.hip) ----
// fdev: the fluid in device memory
void launchBoAssemble(const BoFluid* fdev, const BoGrid& G, const double* pv, const uint8_t* mean,
                      const double* poro, const double* trans, const double* st1, double storefac,
                      double* res, double* jac, const int* diagOff, const int* nbOff, hipStream_t s,
                      const BoDiff* dif = nullptr, const BoDisp* dsp = nullptr, const BoGeo* geo = nullptr);
// the cell-wise source pass over the sources grouped by cell (groups[g] =
// {first, count} into order); vol: nullptr (every cell G.V) or the cells' volumes
void launchBoSources(const BoFluid& f, const BoGrid& G, const double* pv, const uint8_t* mean,
                     int ngroups, const int2* groups, const int* order, const int* cell, const int* type,
                     const double* a, const double* b, double* res, double* jac, const int* diagOff, hipStream_t s,
                     const double* vol = nullptr, const uint8_t* domain = nullptr);
// resetSystem_(domain): zero the residual and Jacobian rows of n cells
void launchClearRows(int n, const int* cells, const int* rowPtr, double* jac, double* res, hipStream_t s);
void launchBoStorage(const BoFluid& f, long N, const double* pv, const uint8_t* mean,
                     const double* poro, double* st, hipStream_t s, const BoRock* rock = nullptr);
void launchBoSourceBlocks(const BoFluid& f, const BoGrid& G, const double* pv, const uint8_t* mean,
                          int nsrc, const int* cell, const int* type, const double* a, const double* b,
                          double* outRes, double* outDiag, hipStream_t s, const double* vol = nullptr);
void launchBoBoundary(const BoFluid* fp, const BoGrid& G, const double* pv, const uint8_t* mean, const BoBc& bc,
                      const double* depth, double* res, double* jac, const int* diagOff, hipStream_t s,
                      double* flows = nullptr);
void launchBoFlows(const BoFluid* fdev, const BoGrid& G, const double* pv, const uint8_t* mean, const double* trans,
                   const int* rowPtr, double* flows, double* flores, hipStream_t s, const BoDiff* dif = nullptr,
                   const BoDisp* dsp = nullptr, const BoGeo* geo = nullptr);
constexpr int kOutFields = 18;
void launchBoOutput(const BoFluid* fdev, long N, const double* pv, const uint8_t* mean, const double* poro,
                    double* out, hipStream_t s, const BoRock* rock = nullptr);
void launchAddSparseSources(int n, const int* cell, const double* r, const double* d, double* res,
                            double* jac, const int* diagOff, hipStream_t s);
void launchBoUpdate(const BoFluid& f, long N, double* pv, uint8_t* mean, uint8_t* wasSw,
                    const double* delta, const NewtonUpdateParams& prm, int* counters, hipStream_t s);

// ---- linear algebra kernels (linalg_kernels.hip) ----
// Scalars of one BiCGStab run kept on the device (no host round trip per dot).
struct KrylovScalars {
    double rho, alpha, omega, beta, rho_i, denom;
    double resnorm, init_norm;
    int fail_code;        // 0 ok, 2 breakdown (NumericalProblem)
    int fail_tag;         // half-iteration index of the first breakdown
    int stag_half;        // change indicator all zero at the last half-step check
    int stag_full_tag;    // half-iteration index of a full-step stagnation, -1 if none
    int tag_base;         // half-iteration index of the current iteration's first half (2 it)
    int pad_;
    // the test after each first half, evaluated on the device so that the next
    // iteration can be queued before the host has seen this one's result
    double tol, abs_tol, max_error;
    int run;              // 1 while the solve goes on; every kernel of the solve returns at once when 0
    int stop_rc;          // the solve's return code once run is 0
    int stop_back;        // iterations the count goes back (0, or 1 for a test of the previous second half)
    int pad2_;
};

constexpr int kRedBlocks = 1024;     // partial sums per reduction (fixed => deterministic)

// run: nullptr, or a device flag; the product is skipped while it is 0
template <class S> void launchSpmv(int n, int bs, const int* rp, const int* ci, const S* a,
                                   const S* x, S* y, hipStream_t st, const int* run = nullptr);
template <class S> void launchCopyConvertD(long m, const double* x, S* y, hipStream_t st);
// the structured layout D[7][N][9] of a 7-point 3x3 BCRS matrix on an
// nx x ny x nz grid (slots -z, -y, -x, I, +x, +y, +z; zero blocks where a
// neighbour is absent), converted to S, and y = A x in it (same bits as launchSpmv)
template <class S>
void launchToDia(int nx, int ny, int nz, const int* rp, const double* A, S* D, hipStream_t st);
template <class S>
void launchSpmvDia(int nx, int ny, int nz, const S* D, const S* x, S* y, hipStream_t st, const int* run = nullptr);

// ILU0 with level scheduling; lvlPtr/lvlRows on the device, lvlPtrHost on the host
struct LevelSchedule {
    int nLevels = 0;
    std::vector<int> ptrHost;     // [nLevels+1]
    DevBuf<int> rows;             // rows ordered by level
};
// the levels of a block pattern's rows for the triangular sweeps: a row's
// level is one more than the largest level of the lower-numbered rows it is
// coupled with in either direction, so the forward sweep (levels ascending)
// and the backward sweep (descending) find every row they read done also
// when the pattern is not symmetric
void buildLevelSchedule(long rows, const std::vector<int>& rowPtr, const std::vector<int>& colIdx,
                        LevelSchedule& L, hipStream_t s);
void launchAddAuxiliary(int nb, const int* blockOff, const double* blocks, int nr, const int* resRow,
                        const double* resVals, double* jac, double* res, hipStream_t s);
void launchAuxUpdate(long n, double* x, const double* delta, int* counters, hipStream_t s);
template <class S> int launchIlu0Factor(int n, int bs, const int* rp, const int* ci, const int* diag,
                                        S* lu, const LevelSchedule& L, int* dFail, hipStream_t st);
template <class S, class P> void launchIlu0Apply(int n, int bs, const int* rp, const int* ci,
                                                 const int* diag, const S* lu, const P* dinv,
                                                 const S* d, S* v, const LevelSchedule& L, hipStream_t st);
template <class S> void launchJacobiInvert(int n, int bs, const int* diag, const S* a, S* dinv,
                                           int* dFail, hipStream_t st);
// `sweeps` block-Jacobi sweeps on v (3x3 blocks) from its current content; tmp: scratch of v's size
template <class S> void launchJacobiSweeps(int n, int sweeps, const int* rp, const int* ci, const S* a,
                                           const double* dinv, const S* d, S* v, S* tmp, double w, const int* run,
                                           hipStream_t st);
template <class S> void launchExtractInvDiag(int n, int bs, const int* diag, const S* lu, double* dinv, hipStream_t st);

// BiCGStab steps (see bicgstab in host_solver.cpp for the sequence)
// tagOff: 0 for the first half of an iteration, 1 for the second (tag = ks->tag_base + tagOff)
// Scratch of the exact scalar products (exact_sum.h): the device
// accumulators of up to two sums and the ticket of the workgroup that
// finishes last (all zero between products: that workgroup clears them).
constexpr int kDotMaxBlocks = 256;
constexpr int kDotThreads = 512;
struct DotScratch {
    ExactAcc acc[2];
    unsigned int ticket;
    unsigned int pad_[63];
};
enum DotMode : int { kDotRho = 0, kDotAlpha = 1, kDotOmega = 2, kDotValue = 3 };
// one launch per product: op 0 rho_i/beta, op 1 denom/alpha
template <class S> void launchDotFinal(long m, const S* x, const S* y, DotScratch* sc, KrylovScalars* ks,
                                       int op, int tagOff, hipStream_t st);
// out = <x, y> on the device
void launchScalarProduct(long m, const double* x, const double* y, DotScratch* sc, double* out, hipStream_t st);
// the same with the accumulator in a static device array (noalloc.hip): no allocation
void scalarProductNoAlloc(long m, const double* x, const double* y, double* out, hipStream_t st);
template <class S> void launchUpdateP(long m, const S* r, const S* v, S* p, const KrylovScalars* ks,
                                      hipStream_t st);
template <class S> void launchUpdateXR(long m, const S* y, const S* v, S* x, S* r, double* partial,
                                       KrylovScalars* ks, hipStream_t st);
// the second half ends the iteration: its final kernel advances tag_base by `advance`
template <class S> void launchUpdateXR2(long m, const S* z, const S* t, S* x, S* r, double* partial,
                                        KrylovScalars* ks, int tagOff, int advance, hipStream_t st);
template <class S> void launchDot2Final(long m, const S* t, const S* s, DotScratch* sc, KrylovScalars* ks,
                                        int tagOff, hipStream_t st);
template <class S> void launchMaxAbs(long m, const S* x, double* partial, double* out, hipStream_t st);
void launchResidualUpdate(int n, int bs, const int* rp, const int* ci, const double* a, const double* x,
                          const double* b, double* r, hipStream_t st);
void launchAxpyD(long m, double alpha, const double* x, double* y, hipStream_t st);
void launchAddFromFloat(long m, const float* d, double* x, hipStream_t st);

}  // namespace omodels
