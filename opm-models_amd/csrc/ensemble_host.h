// ensemble_host.h -- host classes of the ensemble engine and of the serving
// loop (see ensemble.cpp).
#pragma once

#include <memory>
#include <vector>

#include "../../include/omodels.h"
#include "ens_ops.h"
#include "ensemble.h"
#include "omodels_host.h"

namespace omodels {

class Ensemble {
public:
    static constexpr int kMaxMembers = 1024;

    struct Member {
        int id = 0;
        HostTopology topo;
        long N = 0, nnzb = 0;
        BoGrid G{};
        TileGeom tg{};
        int gx = 0, gy = 0, gz = 0, kChunk = 1, nsrc = 0;
        DevBuf<double> trans, poro, pv0, pv1, st1, res, jac, lu64, dinv64, srcA, srcB;
        DevBuf<uint8_t> mean0, mean1, wasSw;
        DevBuf<int> rowPtr, colIdx, diagOff, nbOff, srcCell, srcType, counters, srcOrder;
        DevBuf<int2> srcGroups;
        int nSrcGroups = 0;
        DevBuf<float> Af, luf;
        DevBuf<double> dv[kNumVec];
        DevBuf<float> fv[kNumVec];
        // time stepping and Newton state
        int stepsTotal = 0, stepsDone = 0, attempt = 0, numIter = 0, error = 0;
        int phase = 0;                   // 0: between time steps, 1: in a step's Newton iterations
        double dt = 0.0, errorN = 1e100, lastErrorN = 1e100;
        omodels_run_stats stats{};
        std::vector<omodels_iter_log> log;
        std::vector<double> dtHist;
    };

    explicit Ensemble(const omodels_params& prm, hipStream_t s = nullptr);
    ~Ensemble();
    Ensemble(const Ensemble&) = delete;
    Ensemble& operator=(const Ensemble&) = delete;

    // device bytes one member of this problem's size needs (admission control)
    static size_t memoryEstimate(const omodels_blackoil_problem* p, const omodels_params& prm);
    int add(const omodels_blackoil_problem* p, const double* pv, const uint8_t* mean, int nSteps, double dt0);
    void remove(int id);
    Member* find(int id);
    void download(int id, double* pv, uint8_t* mean);
    // one batched Newton iteration of every member with steps left: a member
    // between time steps starts its next step here, a member whose step ends
    // (converged, or failed for good) advances on its own -- no member waits
    // for another's step to end
    void round();
    // set the tiled ILU's error flag as an exhausted dependency wait does
    // (tests of the recovery; not with the two-colour ILU)
    void raiseIluError();
    bool allDone() const;
    int size() const { return (int)members_.size(); }
    KernelTimer timer;

private:
    void rebuild();
    void refreshStorefac();
    void setFlags(DevBuf<int>& dst, const std::vector<int>& flags);
    std::vector<IluInst> iluInsts(const int* act, bool floatFactors, int idD, int idV) const;
    std::vector<ColInst> colInsts(const int* act, bool floatMatrix, int idD, int idV) const;
    bool colored() const { return prm_.precond == 0 && prm_.ilu_order == 1; }
    void linearSolve(const std::vector<int>& set, std::vector<int>& rc, std::vector<int>& its);
    void linearSolveImpl(const std::vector<int>& set, std::vector<int>& rc, std::vector<int>& its);
    template <class S>
    void bicgstab(const std::vector<int>& set, const std::vector<double>& tol, double absTol, int maxIter,
                  std::vector<int>& rc, std::vector<int>& its);

    omodels_params prm_;
    hipStream_t stream_ = nullptr;
    bool ownStream_ = false;
    BoFluid fluid_{};
    DevBuf<BoFluid> fluidDev_;   // fluid_ in device memory, for the assembly
    bool fluidSet_ = false;
    std::vector<std::unique_ptr<Member>> members_;
    int nextId_ = 0;
    bool dirty_ = true;
    // the batched BiCGStab half-iterations as executable graphs, [FP64, FP32][half];
    // valid until the member set changes
    hipGraphExec_t graphs_[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    void freeGraphs();
    std::unique_ptr<TiledIluSet> ilu_;        // natural order (ilu_order 0)
    std::unique_ptr<ColoredIluSet> col_;      // two-colour tiles (ilu_order 1)
    DevBuf<EnsInst> desc_;
    std::vector<EnsInst> hostDesc_;
    DevBuf<int2> asmTable_, chunkTable_, srcTable_;
    DevBuf<int> first_, actLin_, actRefresh_, actSolve_, actInner_, fail_;
    int nAsm_ = 0, nChunks_ = 0, nSrc_ = 0;
    DevBuf<double> partial_, maxOut_;
    DevBuf<KrylovScalars> ks_;
    DevBuf<int4> dotItems_;          // work items of the scalar products
    DevBuf<ExactAcc> dotAcc_;        // two per member (zero between products)
    DevBuf<unsigned int> tickets_;   // per member (zero between products)
    int nDotItems_ = 0;
    EnsDots dots() const { return EnsDots{dotItems_.p, nDotItems_, dotAcc_.p, tickets_.p}; }
    KrylovScalars* hostKs_ = nullptr;
    double* hostD_ = nullptr;
    int* hostI_ = nullptr;
};

// Request loop with HBM-budget admission and continuous batching.
class Server {
public:
    struct Request {
        int id = 0;
        omodels_blackoil_problem prob{};
        std::vector<double> perm, poro, oil, gas, srcA, srcB, pv, result, dtHist;
        std::vector<int32_t> srcCell, srcType;
        std::vector<uint8_t> mean, resultMean;
        int nSteps = 0, status = -1;
        double dt0 = 0.0, arrival = 0.0, admit = -1.0, finish = -1.0;
        size_t bytes = 0;
        omodels_run_stats stats{};
        std::vector<omodels_iter_log> log;
        omodels_blackoil_problem view() const;
    };

    Server(const omodels_params& prm, double budgetBytes, int maxBatch);
    int submit(const omodels_blackoil_problem* p, const double* pv, const uint8_t* mean, int nSteps, double dt0,
               double arrival);
    void run();
    const Request& request(int id) const { return reqs_.at(id); }
    int numRequests() const { return (int)reqs_.size(); }
    double wall() const { return wall_; }
    long batchSteps() const { return steps_; }
    int maxRunning() const { return maxRunning_; }
    Ensemble& ensemble() { return ens_; }

private:
    omodels_params prm_;
    double budget_;
    int maxBatch_;
    Ensemble ens_;
    std::vector<Request> reqs_;
    double wall_ = 0.0;
    long steps_ = 0;
    int maxRunning_ = 0;
};

}  // namespace omodels
