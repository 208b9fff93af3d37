// omodels_host.h -- host-side C++ classes behind the C ABI.  They mirror the
// reference's interfaces for the hot path:
//   Model                 FvBaseDiscretization state (solution levels, storage cache)
//   TpfaLinearizer        tpfalinearizer.hh:75-951
//   BiCGStabBackend       parallelbicgstabbackend.hh:91-172 / parallelbasebackend.hh:107-386
//   BlackOilNewtonMethod  newtonmethod.hh:90-877 / blackoilnewtonmethod.hh:54-448
#pragma once

#include <array>
#include <memory>
#include <vector>

#include "../../include/omodels.h"
#include "omodels_internal.h"
#include "amg.h"
#include "ilu_colored.h"
#include "ilu_tiled.h"

namespace omodels {

struct HostTopology {
    int nx = 0, ny = 0, nz = 0;
    long N = 0;
    std::vector<int32_t> rowPtr, colIdx, diagOff, nbOff;   // nbOff [6][N], -1 = no neighbour
    std::vector<double> trans;                             // [3][N] faces to +x, +y, +z
    std::vector<int32_t> levelPtr, levelRows;
};

void defaultFluid(std::vector<double>& oil, int& nOil, std::vector<double>& gas, int& nGas, double c[7]);
BoFluid makeFluid(const omodels_blackoil_problem* p);
// a piecewise-linear table of n in [2, kMaxTab] nodes with increasing x
void fillTab(Tab& t, const double* x, const double* y, int n);
void generate(const omodels_blackoil_spec& sp, std::vector<double>& perm, std::vector<double>& poro,
              std::vector<double>& pv, std::vector<uint8_t>& mean, std::vector<int32_t>& srcCell,
              std::vector<int32_t>& srcType, std::vector<double>& srcA, std::vector<double>& srcB);
void buildTopology(const omodels_blackoil_problem* p, HostTopology& T);

// record the work `body` puts on stream s into an executable graph
template <class F>
hipGraphExec_t captureGraph(hipStream_t s, F&& body)
{
    OM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipGraph_t g = nullptr;
    try {
        body();
    } catch (...) {
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    OM_HIP(hipStreamEndCapture(s, &g));
    hipGraphExec_t e = nullptr;
    const hipError_t rc = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    OM_HIP(rc);
    return e;
}
// The problem's sources grouped by cell, for the cell-wise source pass:
// groups[g] = {first, count} into order (source indices in the list's order),
// one group per cell with sources, in the order of each cell's first source.
void sourceGroups(const int32_t* cell, int n, std::vector<int2>& groups, std::vector<int>& order);

// ---------------------------------------------------------------------------
// The linearized system as the linear-solver backend sees it (the
// linearizer's jacobian() and residual(), tpfalinearizer.hh:893-896): block
// CSR pattern, FP64 values and right-hand side on the device, plus the level
// schedule of the pattern and the stream everything runs on.
struct LinearSystem {
    long N = 0, nnzb = 0;
    long nAux = 0;               // auxiliary degrees of freedom after the N grid rows
    long rows() const { return N + nAux; }
    int bs = 3;                  // block size (3 black-oil, 2 two-phase)
    int gx = 0, gy = 0, gz = 0;  // structured grid sizes (tile-wavefront ILU, bs 3)
    hipStream_t stream = nullptr;
    bool ownStream = false;
    DevBuf<double> res, jac;
    DevBuf<int> rowPtr, colIdx, diagOff;
    LevelSchedule levels;
    DevBuf<double> partial;      // reduction partials
    DevBuf<DotScratch> dots;     // scratch of the exact scalar products
    DevBuf<double> scal;         // device scalars (norms)
    double* hostScal = nullptr;  // pinned mirror
    KernelTimer timer;

    LinearSystem() = default;
    LinearSystem(const LinearSystem&) = delete;
    LinearSystem& operator=(const LinearSystem&) = delete;
    virtual ~LinearSystem();
    void sync();
    double maxAbsResidual();

protected:
    void initSystem(hipStream_t s);   // stream and reduction scratch
};

class Model : public LinearSystem {
public:
    explicit Model(const omodels_blackoil_problem* prob, hipStream_t stream = nullptr);

    BoGrid G{};
    BoFluid fluid{};
    HostTopology topo;

    // device state
    DevBuf<double> trans, poro, pv0, pv1, st1;
    DevBuf<BoFluid> fluidDev;    // the fluid for kernels that read it through a pointer
    DevBuf<uint8_t> mean0, mean1, wasSw;
    DevBuf<int> nbOff;
    DevBuf<int> srcCell, srcType;
    DevBuf<double> srcA, srcB;
    int nsrc = 0;
    DevBuf<int2> srcGroups;      // sources grouped by cell (sourceGroups)
    DevBuf<int> srcOrder;
    int nSrcGroups = 0;
    DevBuf<int> counters;        // [0] switched, [1] non-finite update

    // TpfaLinearizer
    void linearizeDomain(double dt, int iterationIndex);
    // linearizeDomain(domain) of a sub-domain (tpfalinearizer.hh:226-245): n
    // distinct cells
    void linearizeSubdomain(double dt, int iterationIndex, int n, const int* cells);
    DevBuf<uint8_t> domMask;
    DevBuf<int> domCells;
    DevBuf<double> ownDepth, ownVol, ownThp;   // the grid's own geometry for the extended form
    // SeparateSparseSourceTerms (tpfalinearizer.hh:56): when set, linearizeDomain
    // leaves the source terms out of the cell loop and adds them after it as
    // sparse blocks
    bool separateSparseSources = false;
    DevBuf<double> srcRes, srcDiag;     // built-in source terms as blocks, [nsrc][3] and [nsrc][9]
    // the built-in source terms at solution(0) as blocks (in srcRes / srcDiag)
    void computeSourceBlocks();
    // wellModel().addReservoirSourceTerms (tpfalinearizer.hh:843-846): residual and
    // diagonal blocks of n cells (device arrays) added in the given order
    void addSparseSources(int n, const int* cells, const double* r, const double* d);
    // the built-in source terms through the sparse path (the built-in well model)
    void addBuiltinSparseSources();
    // updateFlowsInfo (tpfalinearizer.hh:647-705) at solution(0): device arrays
    // [nnzb - N][3] in the pattern's neighbour order, either may be nullptr
    void computeFlows(double* flows, double* flores);
    // BlackOilDiffusionModule in the face term (enableDiffusion); d == nullptr
    // switches it off.  Without face diffusivities they are derived from the
    // porosity like the transmissibilities.
    void setDiffusion(const omodels_diffusion* d);
    DevBuf<double> faceDiff;     // [3][N], at the transmissibility's index
    BoDiff diffusion{};          // diffusion.diff == nullptr: off
    // BlackOilDispersionModule in the face term; d == nullptr switches it off.
    // Switching it on zeroes the cells' oil velocities, which every
    // linearization then replaces for the next one.
    void setDispersion(const omodels_dispersion* d);
    DevBuf<double> faceDisp, nvA, nvB;
    BoDisp dispersion{};         // dispersion.disp == nullptr: off
    // updateStoredTransmissibilities (tpfalinearizer.hh:868-883) and the other
    // per-neighbour inputs of createMatrix_ (:443-452) from the caller:
    // transmissibilities [3][N] (face I|I+e_axis at axis*N + I), depths [N],
    // volumes [N], threshold pressures [3][N]; each nullptr for the grid's own.
    // With depths, volumes or threshold pressures the assembly reads them
    // (BoGeo); without, it is the plain form of the structured grid.
    void updateDiscretization(const double* faceTrans, const double* depth, const double* volume,
                              const double* thpres);
    std::vector<double> transGrid;       // the grid's own transmissibilities
    std::vector<double> depthHost;       // the caller's depths (empty: the grid's own)
    DevBuf<double> geoDepth, geoVol, geoThp;
    bool geoGiven = false;               // depths, volumes or thresholds from the caller
    BoGeo geo{};                         // geo.depth == nullptr: the grid's own, no rock terms
    // rockCompressibility / rockReferencePressure per cell (comp, pref) and the
    // rockCompPoroMultiplier / rockCompTransMultiplier tables of the oil
    // pressure (table [3][nTab]: p, pore-volume and transmissibility
    // multipliers), blackoilintensivequantities.hh:447-480; all null / 0: none
    void setRock(const double* comp, const double* pref, int nTab, const double* table);
    DevBuf<double> rockComp, rockPref;
    DevBuf<Tab> rockTabs;
    // directional relative permeability multipliers [3 axes][3 phases][N]
    void setDirectionalMobility(const double* mult);
    DevBuf<double> dirMob;
    // boundary conditions (see omodels_model_set_boundary_conditions)
    void setBoundary(int n, const int* cell, const int* face, const int* type, const double* rate,
                     const double* state, const uint8_t* meaning, const double* trans, const double* zface);
    BoBc bc{};
    std::vector<int> bcCellHost, bcFaceHost;   // the faces in cell and face order
    int boundaryFlows(double* hostFlows, int* hostCell, int* hostFace);
    DevBuf<int2> bcGroups;
    DevBuf<int> bcCell, bcFace, bcType;
    DevBuf<double> bcRate, bcState, bcTrans, bcZ;
    DevBuf<uint8_t> bcMean;
    std::vector<double> permHost;
    // auxiliary equations (BaseAuxiliaryModule): n degrees of freedom after
    // the cells and the pattern entries (row[k], col[k]) their module adds
    // (addNeighbors); n = 0 returns to the grid's pattern
    void setAuxiliary(int n, int nLinks, const int* row, const int* col);
    // the module's linearize(matrix, residual): blocks (device offsets into
    // the Jacobian, [n][9]) and residual rows ([n][3]) added in order
    void addAuxiliary(int nb, const int* blockOff, const double* blocks, int nr, const int* resRow,
                      const double* resVals);
    // block offset of (row, col) in the current pattern, -1 if absent
    long blockOffset(long row, long col) const;
    std::vector<int> xRowPtr, xColIdx, xDiagOff, xNbOff;   // the pattern with the auxiliary entries
    DevBuf<int> gridRowPtr;      // the grid's own row pointers (FLOWS' neighbour order)
    DevBuf<double> aux0, aux1;   // the auxiliary DOFs' solution(0) and solution(1), [nAux][3]
    // linearizeAuxiliaryEquations of the Newton loop: called after each
    // linearizeDomain (nonzero return: the linearization failed)
    int (*auxFn)(void* user, void* handle, int iteration) = nullptr;
    void* auxUser = nullptr;
    void* auxHandle = nullptr;
    // the kernels' BoGeo from the given geometry and the rock terms
    void refreshGeo();
    void ensureJacobian();
    void eraseMatrix() { jac.release(); }
    // state management (FvBaseDiscretization::advanceTimeLevel / updateFailed)
    void advanceTimeLevel();
    void restoreTimeLevel();
};

struct SolverStats {
    int iterations = 0;
};

class BiCGStabBackend {
public:
    BiCGStabBackend(LinearSystem& m, const omodels_params& prm);
    ~BiCGStabBackend();
    // prepare(J, r) + setResidual(r): uses the model's Jacobian / residual
    void prepare(const double* hostValues = nullptr, const double* hostResidual = nullptr);
    // solve(x); returns 0 converged, 1 not converged, 2 numerical problem
    int solve();
    int iterations() const { return iters_; }
    const double* solution() const { return x64_.p; }
    void downloadIlu(double* hostValues);
    void applyPrecond(const double* hostD, double* hostV);
    // the aggregation AMG (precond = 2) and its level matrices
    const AmgHierarchy& amg() const;
    void downloadAmgLevel(int level, double* hostValues);

private:
    template <class S> int bicgstab(const S* A, const S* lu, const void* dinv, const S* b, S* x,
                                    double tol, double absTol, int maxIter, int* its);
    // v = M^-1 d; with A and Av also Av = A v inside the preconditioner where it
    // can (returns whether it did)
    template <class S> bool precApply(const S* lu, const void* dinv, const S* d, S* v, const int* run = nullptr,
                                      const S* A = nullptr, S* Av = nullptr);
    int solveImpl();
    void tuneIlu();   // pick the tile shape / schedule of the ILU(0) by timing them

    // One BiCGStab iteration as replayed hipGraphs: `first` holds the first
    // half of iteration 0 (up to the convergence test), `iter` the
    // second half of iteration it-1 followed by the first half of iteration
    // it.  Captured for one set of buffers (key) and one generation of the
    // tiled ILU's resident descriptors; recaptured when either changes.
    struct IterGraphs {
        std::array<const void*, 12> key{};
        long generation = -1;
        hipGraphExec_t first = nullptr, iter = nullptr;
        void reset();
        ~IterGraphs() { reset(); }
    };
    IterGraphs graphs_[2];   // FP64, FP32 Krylov vectors

    LinearSystem& m_;
    long builtAux_ = 0, builtNnzb_ = 0;   // the system it was created on (auxiliary DOFs, blocks)
    omodels_params prm_;
    int iters_ = 0;
    const double* A64_ = nullptr;        // FP64 matrix used by the solver
    DevBuf<double> Acopy_;               // host-provided matrix
    DevBuf<double> b64_, x64_, r64_;
    DevBuf<double> lu64_, dinv64_;
    DevBuf<float> Af_, luf_, bf_, xf_;
    // Krylov work vectors (either precision)
    DevBuf<char> jacTmp_;        // block Jacobi: the previous sweep's vector
    DevBuf<double> wd_[5];
    DevBuf<float> wf_[5];
    DevBuf<KrylovScalars> ks_;
    KrylovScalars* ksHost_ = nullptr;    // two pinned copies: iteration it lands in ksHost_[it % 2]
    hipEvent_t ksEv_[2] = {nullptr, nullptr};
    DevBuf<int> fail_;
    std::unique_ptr<TiledIlu> tiled_;
    std::unique_ptr<ColoredIlu> colored_;   // ilu_order 1
    std::unique_ptr<AmgHierarchy> amg_;     // precond 2
    // with the two-colour ILU the solver's matrix in the structured layout
    // (the sweeps and the operator read it; FP64 for the FP64 solve, FP32 for
    // the FP32 inner solve)
    DevBuf<double> dia64_;
    DevBuf<float> diaf_;
    template <class S> const S* diaOf() const
    {
        if constexpr (sizeof(S) == 8) return dia64_.p;
        else return diaf_.p;
    }
    // the inner solve's operator: the structured layout with the two-colour
    // ILU, the BCRS matrix otherwise
    template <class S> void spmvInner(const S* A, const S* x, S* y, const int* run);
    // a 3x3 system on the model's nx x ny x nz grid (7-point pattern)
    bool structured() const;
    bool prepared_ = false;
    bool tuned_ = false;
};

class BlackOilNewtonMethod {
public:
    BlackOilNewtonMethod(Model& m, const omodels_params& prm) : m_(m), prm_(prm), solver_(m, prm) {}
    // NewtonMethod::apply (newtonmethod.hh:228-438); true when converged
    bool apply(double dt, int step, int attempt, std::vector<omodels_iter_log>* log, omodels_run_stats* st);
    int numIterations() const { return numIter_; }
    double suggestTimeStepSize(double oldDt) const;
    int update(const double* delta, int* numSwitched);

private:
    Model& m_;
    omodels_params prm_;
    BiCGStabBackend solver_;
    int numIter_ = 0;
    double error_ = 1e100, lastError_ = 1e100;
};

int simulate(Model& m, const omodels_params& prm, int nSteps, double dt0, double* pvHist, double* dtHist,
             std::vector<omodels_iter_log>* log, omodels_run_stats* stats);

}  // namespace omodels
