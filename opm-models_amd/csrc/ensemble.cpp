// ensemble.cpp -- ensemble engine and serving loop.
//
// An Ensemble holds many black-oil reservoirs ("members") of different sizes,
// each with its own device-resident state (solution levels, storage cache,
// Jacobian, factors, Krylov vectors).  One launch of each kernel covers all
// members (per-instance descriptors and block tables), while the control flow
// of every member -- Newton convergence, linear-solver termination, time-step
// halving on failure and the next step size -- is decided per member exactly
// as NewtonMethod::apply / FvBaseProblem::timeIntegration do for one
// reservoir (newtonmethod.hh:228-465, fvbaseproblem.hh:512-606).
//
// The Server admits requests (a reservoir plus a number of time steps) in
// arrival order whenever their estimated device memory fits the HBM budget,
// runs one lockstep time step of all admitted members at a time, and retires
// members as soon as their last step is done, so requests that arrive while
// others run join the next step (continuous batching).
#include <algorithm>
#include <cstdio>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>

#include "ens_ops.h"
#include "ensemble_host.h"
#include "omodels_host.h"

namespace omodels {

// ---------------------------------------------------------------------------
size_t Ensemble::memoryEstimate(const omodels_blackoil_problem* p, const omodels_params& prm)
{
    const size_t N = (size_t)p->nx * p->ny * p->nz;
    const size_t nnzb = 7 * N;
    const int precision = prm.precision;
    const bool colored = prm.precond == 0 && prm.ilu_order == 1;
    // J, the ILU factor (the two-colour ILU keeps only pivots), FP64 pivots
    size_t b = nnzb * 9 * 8 * (colored ? 1 : 2) + N * 9 * 8;
    if (precision == 1) b += nnzb * 9 * 4 * (colored ? 1 : 2);   // FP32 J and factor
    b += N * 3 * 8 * (kNumVec + 5) + N * 3 * 4 * kNumVec;      // vectors and state
    b += N * (4 * 8 + 6 * 4 + 3);                            // topology, transmissibilities, flags
    if (!colored) b += N * 9 * 8 + (precision == 1 ? N * 9 * 4 : 0);   // exchange slots of the tiled sweeps
    return b;
}

Ensemble::Ensemble(const omodels_params& prm, hipStream_t s) : prm_(prm)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("no HIP device available");
    if (s) stream_ = s;
    else { OM_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking)); ownStream_ = true; }
    ilu_ = std::make_unique<TiledIluSet>(stream_);
    // members always use the tiled kernels, by default in the batched sweep form:
    // an ensemble is throughput-bound, and the batched form does less work per
    // tile (serving stream of 32 requests: 1.57 solves/s batched, 1.48 with the
    // form chosen per launch, 1.42 streaming); ilu_schedule 2 forces streaming
    ilu_->setSweepMode(prm_.ilu_schedule == 2 ? 1 : 2);
    if (prm_.ilu_order < 0 || prm_.ilu_order > 1) throw std::invalid_argument("ilu_order must be 0 or 1");
    if (colored()) col_ = std::make_unique<ColoredIluSet>();
    OM_HIP(hipHostMalloc((void**)&hostKs_, sizeof(KrylovScalars) * kMaxMembers, hipHostMallocDefault));
    OM_HIP(hipHostMalloc((void**)&hostD_, sizeof(double) * kMaxMembers, hipHostMallocDefault));
    OM_HIP(hipHostMalloc((void**)&hostI_, sizeof(int) * 2 * kMaxMembers, hipHostMallocDefault));
    ks_.alloc(kMaxMembers);
    tickets_.alloc(kMaxMembers);
    OM_HIP(hipMemsetAsync(tickets_.p, 0, kMaxMembers * sizeof(unsigned int), stream_));
    dotAcc_.alloc(2 * kMaxMembers);
    OM_HIP(hipMemsetAsync(dotAcc_.p, 0, 2 * kMaxMembers * sizeof(ExactAcc), stream_));
    maxOut_.alloc(kMaxMembers);
    for (auto* a : {&actLin_, &actRefresh_, &actSolve_, &actInner_}) a->alloc(kMaxMembers);
    fail_.alloc(1);
}

void Ensemble::freeGraphs()
{
    for (auto& g : graphs_)
        for (auto& h : g) {
            if (h) (void)hipGraphExecDestroy(h);
            h = nullptr;
        }
}

Ensemble::~Ensemble()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    freeGraphs();
    members_.clear();
    if (hostKs_) (void)hipHostFree(hostKs_);
    if (hostD_) (void)hipHostFree(hostD_);
    if (hostI_) (void)hipHostFree(hostI_);
    if (ownStream_ && stream_) (void)hipStreamDestroy(stream_);
}

int Ensemble::add(const omodels_blackoil_problem* p, const double* pv, const uint8_t* mean, int nSteps, double dt0)
{
    if ((int)members_.size() >= kMaxMembers) throw std::invalid_argument("ensemble is full");
    if (nSteps < 1 || !(dt0 > 0)) throw std::invalid_argument("invalid step count or dt");
    BoFluid f = makeFluid(p);
    if (members_.empty() && !fluidSet_) {
        fluid_ = f;
        fluidSet_ = true;
        fluidDev_.upload(&fluid_, 1, stream_);
    }
    else if (std::memcmp(&f, &fluid_, sizeof(BoFluid)) != 0)
        throw std::invalid_argument("all members of an ensemble share one fluid description");
    auto m = std::make_unique<Member>();
    Member& M = *m;
    M.id = nextId_++;
    buildTopology(p, M.topo);
    M.N = M.topo.N;
    M.nnzb = M.topo.rowPtr[M.N];
    M.G.nx = p->nx; M.G.ny = p->ny; M.G.nz = p->nz;
    M.G.dx = p->dx; M.G.dy = p->dy; M.G.dz = p->dz; M.G.g = p->gravity;
    M.G.V = p->dx * p->dy * p->dz;
    M.G.area[0] = p->dy * p->dz; M.G.area[1] = p->dx * p->dz; M.G.area[2] = p->dx * p->dy;
    TunedIlu tuned;
    if (prm_.autotune && lookupTunedIlu(p->nx, p->ny, p->nz, prm_.precision, tuned) && !tuned.level)
        M.tg = tuned.shape;
    else M.tg = makeTileGeom(p->nx, p->ny, p->nz);
    assemblyGrid(M.G, M.gx, M.gy, M.gz, M.kChunk, 512);
    hipStream_t s = stream_;
    const long N = M.N;
    M.trans.upload(M.topo.trans.data(), M.topo.trans.size(), s);
    M.poro.upload(p->poro, N, s);
    M.rowPtr.upload(M.topo.rowPtr.data(), M.topo.rowPtr.size(), s);
    M.colIdx.upload(M.topo.colIdx.data(), M.topo.colIdx.size(), s);
    M.diagOff.upload(M.topo.diagOff.data(), M.topo.diagOff.size(), s);
    M.nbOff.upload(M.topo.nbOff.data(), M.topo.nbOff.size(), s);
    M.nsrc = p->n_src;
    if (M.nsrc > 0) {
        for (int k = 0; k < M.nsrc; ++k)
            if (p->src_cell[k] < 0 || p->src_cell[k] >= N) throw std::invalid_argument("source cell out of range");
        M.srcCell.upload(p->src_cell, M.nsrc, s);
        M.srcType.upload(p->src_type, M.nsrc, s);
        M.srcA.upload(p->src_a, M.nsrc, s);
        M.srcB.upload(p->src_b, M.nsrc, s);
        std::vector<int2> groups;
        std::vector<int> order;
        sourceGroups(p->src_cell, M.nsrc, groups, order);
        M.nSrcGroups = (int)groups.size();
        M.srcGroups.upload(groups.data(), groups.size(), s);
        M.srcOrder.upload(order.data(), order.size(), s);
        OM_HIP(hipStreamSynchronize(s));
    }
    M.pv0.upload(pv, 3 * N, s);
    M.mean0.upload(mean, N, s);
    M.pv1.upload(pv, 3 * N, s);
    M.mean1.upload(mean, N, s);
    M.wasSw.alloc(N);
    OM_HIP(hipMemsetAsync(M.wasSw.p, 0, N, s));
    M.st1.alloc(3 * N);
    M.res.alloc(3 * N);
    M.jac.alloc(9 * (size_t)M.nnzb);
    if (!colored()) M.lu64.alloc(9 * (size_t)M.nnzb);      // the two-colour ILU keeps only pivots
    M.dinv64.alloc(9 * (size_t)N);
    for (int v = 0; v < kNumVec; ++v) if (v != kVb) M.dv[v].alloc(3 * N);
    if (prm_.precision == 1) {
        // the two-colour ILU reads the FP32 matrix in the structured layout
        M.Af.alloc(9 * (colored() ? 7 * (size_t)M.N : (size_t)M.nnzb));
        if (!colored()) M.luf.alloc(9 * (size_t)M.nnzb);
        for (int v = 0; v < kNumVec; ++v) M.fv[v].alloc(3 * N);
    }
    M.counters.alloc(4);
    M.stepsTotal = nSteps;
    M.dt = dt0;
    OM_HIP(hipStreamSynchronize(s));
    members_.push_back(std::move(m));
    dirty_ = true;
    return members_.back()->id;
}

Ensemble::Member* Ensemble::find(int id)
{
    for (auto& m : members_) if (m->id == id) return m.get();
    return nullptr;
}

void Ensemble::remove(int id)
{
    OM_HIP(hipStreamSynchronize(stream_));
    members_.erase(std::remove_if(members_.begin(), members_.end(),
                                  [&](const std::unique_ptr<Member>& m) { return m->id == id; }),
                   members_.end());
    dirty_ = true;
}

void Ensemble::download(int id, double* pv, uint8_t* mean)
{
    Member* m = find(id);
    if (!m) throw std::invalid_argument("unknown member");
    if (pv) m->pv0.download(pv, 3 * m->N, stream_);
    if (mean) m->mean0.download(mean, m->N, stream_);
    OM_HIP(hipStreamSynchronize(stream_));
}

// tables and descriptors for the current member set and time-step sizes
void Ensemble::rebuild()
{
    OM_HIP(hipStreamSynchronize(stream_));
    freeGraphs();       // they hold the old set's pointers
    const int B = (int)members_.size();
    std::vector<EnsInst> d(B);
    std::vector<int2> asmT, chunkT, srcT;
    std::vector<int> first(B + 1, 0);
    std::vector<TileGeom> geoms(B);
    for (int b = 0; b < B; ++b) {
        Member& M = *members_[b];
        EnsInst& I = d[b];
        std::memset(&I, 0, sizeof(I));
        I.G = M.G; I.N = (int)M.N; I.nnzb = (int)M.nnzb; I.nsrc = M.nsrc; I.kChunk = M.kChunk;
        I.afDia = colored() ? 1 : 0;
        I.gx = M.gx; I.gy = M.gy; I.gz = M.gz; I.nAsm = M.gx * M.gy * M.gz;
        I.storefac = M.G.V / M.dt;
        I.trans = M.trans.p; I.poro = M.poro.p;
        I.pv0 = M.pv0.p; I.mean0 = M.mean0.p; I.wasSw = M.wasSw.p; I.st1 = M.st1.p; I.res = M.res.p;
        I.jac = M.jac.p; I.rowPtr = M.rowPtr.p; I.colIdx = M.colIdx.p; I.diagOff = M.diagOff.p; I.nbOff = M.nbOff.p;
        I.srcCell = M.srcCell.p; I.srcType = M.srcType.p; I.srcA = M.srcA.p; I.srcB = M.srcB.p;
        I.srcGroups = M.srcGroups.p; I.srcOrder = M.srcOrder.p;
        I.lu64 = M.lu64.p; I.dinv64 = M.dinv64.p; I.Af = M.Af.p; I.luf = M.luf.p;
        for (int v = 0; v < kNumVec; ++v) { I.dv[v] = M.dv[v].p; I.fv[v] = M.fv[v].p; }
        I.dv[kVb] = M.res.p;            // the right-hand side is the residual itself
        I.counters = M.counters.p;
        for (int k = 0; k < I.nAsm; ++k) asmT.push_back(make_int2(b, k));
        const int nc = (int)((M.N + kEnsChunk - 1) / kEnsChunk);
        for (int k = 0; k < nc; ++k) chunkT.push_back(make_int2(b, k));
        first[b + 1] = first[b] + nc;
        for (int k = 0; k < M.nSrcGroups; ++k) srcT.push_back(make_int2(b, k));
        geoms[b] = M.tg;
    }
    desc_.upload(d.data(), d.size(), stream_);
    asmTable_.upload(asmT.data(), asmT.size(), stream_);
    chunkTable_.upload(chunkT.data(), chunkT.size(), stream_);
    srcTable_.upload(srcT.data(), srcT.size(), stream_);
    first_.upload(first.data(), first.size(), stream_);
    nAsm_ = (int)asmT.size();
    nChunks_ = (int)chunkT.size();
    nSrc_ = (int)srcT.size();
    partial_.alloc(std::max<size_t>(2 * (size_t)nChunks_, 1));
    std::vector<int4> items;
    for (int b = 0; b < B; ++b) {
        const int nb = (int)((members_[b]->N + kEnsDotCells - 1) / kEnsDotCells);
        for (int k = 0; k < nb; ++k) items.push_back(make_int4(b, k, nb, 0));
    }
    nDotItems_ = (int)items.size();
    dotItems_.upload(items.data(), items.size(), stream_);

    if (colored()) {
        std::vector<ColGeom> cg(B);
        for (int b = 0; b < B; ++b) cg[b] = colGeomOf(members_[b]->G.nx, members_[b]->G.ny, members_[b]->G.nz);
        col_->setInstances(cg, stream_);
    } else {
        ilu_->setInstances(geoms, stream_);
    }
    hostDesc_ = d;
    OM_HIP(hipStreamSynchronize(stream_));
    dirty_ = false;
}

void Ensemble::refreshStorefac()
{
    bool changed = false;
    for (size_t b = 0; b < members_.size(); ++b) {
        const double sf = members_[b]->G.V / members_[b]->dt;
        if (hostDesc_[b].storefac != sf) { hostDesc_[b].storefac = sf; changed = true; }
    }
    if (changed) {
        OM_HIP(hipStreamSynchronize(stream_));
        desc_.upload(hostDesc_.data(), hostDesc_.size(), stream_);
    }
}

void Ensemble::setFlags(DevBuf<int>& dst, const std::vector<int>& flags)
{
    std::copy(flags.begin(), flags.end(), hostI_);
    OM_HIP(hipMemcpyAsync(dst.p, hostI_, flags.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
    OM_HIP(hipStreamSynchronize(stream_));
}

std::vector<IluInst> Ensemble::iluInsts(const int* act, bool floatFactors, int idD, int idV) const
{
    std::vector<IluInst> v(members_.size());
    for (size_t b = 0; b < members_.size(); ++b) {
        const Member& M = *members_[b];
        IluInst& I = v[b];
        std::memset(&I, 0, sizeof(I));
        I.g = M.tg; I.rp = M.rowPtr.p; I.diag = M.diagOff.p;
        I.luw = M.lu64.p; I.dinvw = M.dinv64.p; I.dinv = M.dinv64.p;
        I.active = act + b;
        if (floatFactors) {
            I.lu = M.luf.p;
            I.d = idD == kVb ? (const void*)M.fv[kVb].p : (const void*)M.fv[idD].p;
            I.v = M.fv[idV].p;
        } else {
            I.lu = M.lu64.p;
            I.d = idD == kVb ? (const void*)M.res.p : (const void*)M.dv[idD].p;
            I.v = M.dv[idV].p;
        }
    }
    return v;
}

std::vector<ColInst> Ensemble::colInsts(const int* act, bool floatMatrix, int idD, int idV) const
{
    std::vector<ColInst> v(members_.size());
    for (size_t b = 0; b < members_.size(); ++b) {
        const Member& M = *members_[b];
        ColInst& I = v[b];
        std::memset(&I, 0, sizeof(I));
        I.g = colGeomOf(M.G.nx, M.G.ny, M.G.nz);
        I.rp = M.rowPtr.p;
        I.dinv = M.dinv64.p;
        I.active = act + b;
        if (floatMatrix) {
            I.rp = nullptr;             // Af in the structured layout
            I.A = M.Af.p;
            I.d = M.fv[idD].p;
            I.v = M.fv[idV].p;
        } else {
            I.A = M.jac.p;
            I.d = idD == kVb ? (const void*)M.res.p : (const void*)M.dv[idD].p;
            I.v = M.dv[idV].p;
        }
    }
    return v;
}

// Batched BiCGStabSolver::apply with CombinedCriterion per member.
template <class S>
void Ensemble::bicgstab(const std::vector<int>& set, const std::vector<double>& tol, double absTol, int maxIter,
                        std::vector<int>& rc, std::vector<int>& its)
{
    const int B = (int)members_.size();
    EnsOps ops(desc_.p, chunkTable_.p, nChunks_, first_.p, B, partial_.p, dots(), ks_.p, stream_);
    const bool fl = sizeof(S) == 4;
    std::vector<int> act(set);
    rc.assign(B, 0);
    its.assign(B, 0);
    setFlags(actInner_, act);
    const int* a = actInner_.p;
    ops.copy<S>(a, kVx, -1);
    ops.copy<S>(a, kVr, kVb);
    ops.maxAbs<S>(a, kVr, 0, maxOut_.p);
    OM_HIP(hipMemcpyAsync(hostD_, maxOut_.p, B * sizeof(double), hipMemcpyDeviceToHost, stream_));
    OM_HIP(hipStreamSynchronize(stream_));
    const double minS = (double)std::numeric_limits<S>::min() * 1e10;
    std::vector<double> init(B, 0.0);
    bool any = false;
    for (int b = 0; b < B; ++b) {
        if (!act[b]) continue;
        init[b] = std::max(hostD_[b], minS);
        if (init[b] / std::max(1e-20, init[b]) <= tol[b] || init[b] <= absTol) { act[b] = 0; rc[b] = 0; }
        else any = true;
    }
    if (!any) return;
    for (int b = 0; b < B; ++b) {
        KrylovScalars& k = hostKs_[b];
        std::memset(&k, 0, sizeof(k));
        k.rho = 1; k.alpha = 1; k.omega = 1; k.init_norm = init[b];
        k.fail_tag = -1; k.stag_full_tag = -1;
    }
    OM_HIP(hipMemcpyAsync(ks_.p, hostKs_, B * sizeof(KrylovScalars), hipMemcpyHostToDevice, stream_));
    // the last workgroup of every scalar product clears its member's ticket
    // and accumulators; this clears what a solve interrupted by an error may
    // have left behind
    OM_HIP(hipMemsetAsync(tickets_.p, 0, (size_t)B * sizeof(unsigned int), stream_));
    OM_HIP(hipMemsetAsync(dotAcc_.p, 0, 2 * (size_t)B * sizeof(ExactAcc), stream_));
    setFlags(actInner_, act);
    ops.copy<S>(a, kVv, -1);
    ops.copy<S>(a, kVp, -1);
    const auto pInsts = colored() ? std::vector<IluInst>() : iluInsts(a, fl, kVp, kVy);
    const auto rInsts = colored() ? std::vector<IluInst>() : iluInsts(a, fl, kVr, kVz);
    const auto pCol = colored() ? colInsts(a, fl, kVp, kVy) : std::vector<ColInst>();
    const auto rCol = colored() ? colInsts(a, fl, kVr, kVz) : std::vector<ColInst>();
    // first half: rho, p, y = M^-1 p, v = A y, alpha, x/r update; second
    // half: z = M^-1 r, y = A z, omega, x/r update and the next iteration's
    // tag base (device scalars: no argument changes from one iteration to
    // the next, so either half can be replayed from a graph)
    auto firstHalf = [&]() {
        ops.dot<S>(a, kVb, kVr, 0, 0);
        ops.updateP<S>(a);
        timer.begin(kClsIluApply, stream_);
        if (colored()) col_->apply<S>(pCol, 0, stream_);
        else ilu_->apply<S>(pInsts, stream_);
        timer.end(stream_);
        ops.spmv<S>(a, kVy, kVv);
        ops.dot<S>(a, kVb, kVv, 1, 0);
        ops.updateXR<S>(a);
    };
    auto secondHalf = [&]() {
        timer.begin(kClsIluApply, stream_);
        if (colored()) col_->apply<S>(rCol, 1, stream_);
        else ilu_->apply<S>(rInsts, stream_);
        timer.end(stream_);
        ops.spmv<S>(a, kVz, kVy);
        ops.dot2<S>(a, 1);
        ops.updateXR2<S>(a, 1);
    };
    // graph replay (params.graph) with the two-colour ILU, whose kernels and
    // descriptors are all a graph needs: captured in the first solve of a
    // member set that iterates twice, replayed by every later iteration until
    // the set changes
    const bool useGraph = prm_.graph && colored();
    hipGraphExec_t* gg = graphs_[sizeof(S) == 4 ? 1 : 0];
    if (useGraph) {
        col_->prime(pCol, 0, stream_);
        col_->prime(rCol, 1, stream_);
    }
    // before the first capture, iteration 0 runs eagerly: its kernels set up
    // what they allocate on first use, which a capture must not do
    auto runHalf = [&](int half, int it) {
        if (!useGraph || (!gg[0] && it == 0)) { half == 0 ? firstHalf() : secondHalf(); return; }
        if (!gg[0]) {
            const bool timed = timer.enabled;
            timer.enabled = false;      // event records stay out of the graphs
            try {
                gg[0] = captureGraph(stream_, firstHalf);
                gg[1] = captureGraph(stream_, secondHalf);
            } catch (...) {
                timer.enabled = timed;
                freeGraphs();
                throw;
            }
            timer.enabled = timed;
        }
        OM_HIP(hipGraphLaunch(gg[half], stream_));
    };
    if (prm_.precond != 0) throw std::invalid_argument("the ensemble engine uses the block-ILU0 preconditioner");
    for (int it = 0; it < maxIter; ++it) {
        const int tag = 2 * it;
        runHalf(0, it);
        OM_HIP(hipMemcpyAsync(hostKs_, ks_.p, B * sizeof(KrylovScalars), hipMemcpyDeviceToHost, stream_));
        OM_HIP(hipStreamSynchronize(stream_));
        bool changed = false;
        any = false;
        for (int b = 0; b < B; ++b) {
            if (!act[b]) continue;
            const KrylovScalars& K = hostKs_[b];
            int r = -1;
            if (it > 0 && K.fail_code && K.fail_tag == tag - 1) { r = 2; its[b] = it - 1; }
            else if (it > 0 && K.stag_full_tag == tag - 1) { r = 1; its[b] = it - 1; }
            else if (K.fail_code) { r = 2; its[b] = it; }
            else {
                its[b] = it;
                if (K.resnorm / std::max(1e-20, init[b]) <= tol[b] || K.resnorm <= absTol) r = 0;
                else if (K.stag_half || K.resnorm > prm_.lin_max_error) r = 1;
            }
            if (r >= 0) { rc[b] = r; act[b] = 0; changed = true; }
            else any = true;
        }
        if (changed) setFlags(actInner_, act);
        if (!any) return;
        runHalf(1, it);
    }
    OM_HIP(hipMemcpyAsync(hostKs_, ks_.p, B * sizeof(KrylovScalars), hipMemcpyDeviceToHost, stream_));
    OM_HIP(hipStreamSynchronize(stream_));
    for (int b = 0; b < B; ++b)
        if (act[b]) { its[b] = maxIter; rc[b] = hostKs_[b].fail_code ? 2 : 1; }
}

// prepare + solve for the members in `set`; rc per member (0, 1, 2)
void Ensemble::linearSolve(const std::vector<int>& set, std::vector<int>& rc, std::vector<int>& its)
{
    linearSolveImpl(set, rc, its);
    // A dependency wait of the tiled ILU that ran into its bound (in the
    // factorization or a sweep) leaves results that cannot be trusted.  Its
    // flag is one for all members, so every member of this solve fails it as
    // a NumericalProblem (its step is halved and retried); the flag and the
    // exchange slots are reset for the next solve.
    if (!colored() && ilu_->takeError(stream_)) {
        rc.assign(set.size(), 0);
        for (size_t b = 0; b < set.size(); ++b) if (set[b]) rc[b] = 2;
    }
}

void Ensemble::raiseIluError()
{
    if (colored()) throw std::invalid_argument("the two-colour ILU has no dependency waits");
    ilu_->raiseError(stream_);
}

void Ensemble::linearSolveImpl(const std::vector<int>& set, std::vector<int>& rc, std::vector<int>& its)
{
    const int B = (int)members_.size();
    EnsOps ops(desc_.p, chunkTable_.p, nChunks_, first_.p, B, partial_.p, dots(), ks_.p, stream_);
    setFlags(actSolve_, set);
    const int* a = actSolve_.p;
    // preconditioner: FP64 ILU(0) of the FP64 Jacobian
    for (auto& m : members_) OM_HIP(hipMemsetAsync(m->counters.p, 0, 4 * sizeof(int), stream_));
    int* fail = fail_.p;
    OM_HIP(hipMemsetAsync(fail, 0, sizeof(int), stream_));
    timer.begin(kClsIluFactor, stream_);
    if (colored()) {
        col_->factor(colInsts(a, false, kVr, kVz), fail, stream_);
    } else {
        ops.copyMatrix(a, 0);
        ilu_->factor(iluInsts(a, false, kVr, kVz), fail, stream_);
    }
    timer.end(stream_);
    if (prm_.precision == 1) {
        ops.copyMatrix(a, 1);
        if (!colored()) ops.factorToFloat(a);
    }
    int f = 0;
    OM_HIP(hipMemcpyAsync(&f, fail, sizeof(int), hipMemcpyDeviceToHost, stream_));
    OM_HIP(hipStreamSynchronize(stream_));
    if (f) {
        // a singular pivot somewhere: fall back to failing every member of this solve
        rc.assign(B, 0);
        its.assign(B, 0);
        for (int b = 0; b < B; ++b) if (set[b]) rc[b] = 2;
        return;
    }
    const double absTol = prm_.lin_abs_tol < 0 ? prm_.newton_tol / 100.0 : prm_.lin_abs_tol;
    if (prm_.precision == 0) {
        std::vector<double> tol(B, prm_.lin_tol);
        bicgstab<double>(set, tol, absTol, prm_.lin_max_iter, rc, its);
        return;
    }
    // FP32 inner BiCGStab inside FP64 iterative refinement, per member
    std::vector<int> act(set);
    rc.assign(B, 0);
    its.assign(B, 0);
    ops.copy<double>(a, kVx, -1);
    ops.copy<double>(a, kVr64, kVb);
    ops.maxAbs<double>(a, kVr64, 0, maxOut_.p);
    OM_HIP(hipMemcpyAsync(hostD_, maxOut_.p, B * sizeof(double), hipMemcpyDeviceToHost, stream_));
    OM_HIP(hipStreamSynchronize(stream_));
    std::vector<double> init(B, 1.0), res(B, 1.0);
    for (int b = 0; b < B; ++b) {
        if (!act[b]) continue;
        init[b] = std::max(hostD_[b], std::numeric_limits<double>::min() * 1e10);
        res[b] = init[b];
        if (res[b] / std::max(1e-20, init[b]) <= prm_.lin_tol || res[b] <= absTol) act[b] = 0;
    }
    for (int outer = 0; outer < 8; ++outer) {
        bool any = false;
        for (int b = 0; b < B; ++b) any = any || act[b];
        if (!any) return;
        setFlags(actSolve_, act);
        ops.rhsToFloat(a);
        std::vector<double> tol(B, 0.5);
        for (int b = 0; b < B; ++b) if (act[b]) tol[b] = std::min(0.5, std::max(1e-6, prm_.lin_tol * init[b] / res[b]));
        std::vector<int> irc, iits;
        bicgstab<float>(act, tol, absTol, prm_.lin_max_iter, irc, iits);
        for (int b = 0; b < B; ++b) {
            if (!act[b]) continue;
            its[b] += iits[b];
            if (irc[b] == 2) { rc[b] = 2; act[b] = 0; }
        }
        setFlags(actSolve_, act);
        ops.addCorrection(a);
        ops.residual64(a);
        ops.maxAbs<double>(a, kVr64, 0, maxOut_.p);
        OM_HIP(hipMemcpyAsync(hostD_, maxOut_.p, B * sizeof(double), hipMemcpyDeviceToHost, stream_));
        OM_HIP(hipStreamSynchronize(stream_));
        for (int b = 0; b < B; ++b) {
            if (!act[b]) continue;
            res[b] = hostD_[b];
            if (!std::isfinite(res[b])) { rc[b] = 2; act[b] = 0; }
            else if (res[b] / std::max(1e-20, init[b]) <= prm_.lin_tol || res[b] <= absTol) { rc[b] = 0; act[b] = 0; }
            else if (irc[b] == 1 || res[b] > prm_.lin_max_error || its[b] >= prm_.lin_max_iter) { rc[b] = 1; act[b] = 0; }
        }
    }
    for (int b = 0; b < B; ++b) if (act[b]) rc[b] = 1;
}

// One lockstep time step of every member that still has steps to do.
void Ensemble::round()
{
    if (dirty_) rebuild();
    const int B = (int)members_.size();
    if (B == 0) return;
    using clk = std::chrono::steady_clock;
    enum { kIdle = 0, kNewton = 1, kOk = 2, kFailedStep = 3 };
    std::vector<int> state(B, kIdle);
    for (int b = 0; b < B; ++b) {
        Member& M = *members_[b];
        if (M.stepsDone >= M.stepsTotal || M.error) { M.phase = 0; continue; }
        if (M.phase == 0) {     // the next time step begins
            M.phase = 1; M.attempt = 0; M.numIter = 0; M.errorN = 1e100; M.lastErrorN = 1e100;
        }
        state[b] = kNewton;
    }
    auto restartOrFail = [&](int b) {
        Member& M = *members_[b];
        M.stats.failed_attempts++;
        // updateFailed: solution(0) = solution(1), then halve dt
        OM_HIP(hipMemcpyAsync(M.pv0.p, M.pv1.p, 3 * M.N * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        OM_HIP(hipMemcpyAsync(M.mean0.p, M.mean1.p, M.N, hipMemcpyDeviceToDevice, stream_));
        M.attempt++;
        const double ndt = M.dt / 2.0;
        if (M.attempt >= prm_.max_dt_divisions || M.dt < prm_.min_dt * (1 + 1e-9)) {
            M.error = 3;
            state[b] = kFailedStep;
            return;
        }
        M.dt = ndt < prm_.min_dt ? prm_.min_dt : ndt;
        M.numIter = 0; M.errorN = 1e100; M.lastErrorN = 1e100;
        // the retried attempt linearizes in the next round
    };
    const int tgt = prm_.newton_target_iter;
    EnsOps ops(desc_.p, chunkTable_.p, nChunks_, first_.p, B, partial_.p, dots(), ks_.p, stream_);
    do {
        std::vector<int> lin(B, 0), refresh(B, 0);
        bool any = false;
        for (int b = 0; b < B; ++b) if (state[b] == kNewton) { lin[b] = 1; refresh[b] = members_[b]->numIter == 0; any = true; }
        if (!any) break;
        refreshStorefac();
        auto t0 = clk::now();
        setFlags(actLin_, lin);
        setFlags(actRefresh_, refresh);
        launchEnsStorage(fluid_, desc_.p, chunkTable_.p, nChunks_, actRefresh_.p, stream_);
        timer.begin(kClsAssemble, stream_);
        launchEnsAssemble(fluid_, fluidDev_.p, desc_.p, asmTable_.p, nAsm_, actLin_.p, srcTable_.p, nSrc_, stream_);
        timer.end(stream_);
        ops.maxAbs<double>(actLin_.p, 0, 1, maxOut_.p);
        OM_HIP(hipMemcpyAsync(hostD_, maxOut_.p, B * sizeof(double), hipMemcpyDeviceToHost, stream_));
        OM_HIP(hipStreamSynchronize(stream_));
        auto t1 = clk::now();
        std::vector<int> solve(B, 0);
        bool anySolve = false;
        for (int b = 0; b < B; ++b) {
            if (!lin[b]) continue;
            Member& M = *members_[b];
            M.stats.linearizations++;
            M.stats.t_linearize += std::chrono::duration<double>(t1 - t0).count();
            M.lastErrorN = M.errorN;
            M.errorN = hostD_[b];
            if (!(M.errorN <= prm_.newton_max_error)) { restartOrFail(b); continue; }
            const bool conv = M.errorN <= prm_.newton_tol;
            bool proceed;
            if (M.numIter < 1) proceed = true;
            else if (conv) proceed = false;
            else if (M.numIter >= prm_.newton_max_iter) proceed = M.errorN * 4.0 < M.lastErrorN;
            else proceed = true;
            if (!proceed) {
                M.log.push_back({M.stepsDone, M.attempt, M.numIter, 0, 0, 1, M.errorN, M.dt});
                M.numIter++;
                if (conv) state[b] = kOk;
                else { M.numIter = 2 * tgt; restartOrFail(b); }
                continue;
            }
            solve[b] = 1;
            anySolve = true;
        }
        if (!anySolve) break;
        std::vector<int> rc, its;
        linearSolve(solve, rc, its);
        auto t2 = clk::now();
        std::vector<int> upd(B, 0);
        bool anyUpd = false;
        for (int b = 0; b < B; ++b) {
            if (!solve[b]) continue;
            Member& M = *members_[b];
            M.stats.t_solve += std::chrono::duration<double>(t2 - t1).count();
            M.stats.linear_iters += its[b];
            if (rc[b] != 0) {
                M.log.push_back({M.stepsDone, M.attempt, M.numIter, its[b], 0, 0, M.errorN, M.dt});
                M.numIter = 2 * tgt;
                restartOrFail(b);
                continue;
            }
            upd[b] = 1;
            anyUpd = true;
        }
        if (!anyUpd) break;
        setFlags(actSolve_, upd);
        for (auto& m : members_) OM_HIP(hipMemsetAsync(m->counters.p, 0, 4 * sizeof(int), stream_));
        NewtonUpdateParams up{prm_.dp_max_rel, prm_.ds_max, prm_.privar_osc_thresh};
        launchEnsUpdate(fluid_, desc_.p, chunkTable_.p, nChunks_, actSolve_.p, up, stream_);
        for (int b = 0; b < B; ++b)
            if (upd[b]) OM_HIP(hipMemcpyAsync(hostI_ + 2 * b, members_[b]->counters.p, 2 * sizeof(int),
                                              hipMemcpyDeviceToHost, stream_));
        OM_HIP(hipStreamSynchronize(stream_));
        auto t3 = clk::now();
        for (int b = 0; b < B; ++b) {
            if (!upd[b]) continue;
            Member& M = *members_[b];
            M.stats.t_update += std::chrono::duration<double>(t3 - t2).count();
            M.stats.newton_iters++;
            M.log.push_back({M.stepsDone, M.attempt, M.numIter, its[b], hostI_[2 * b], 1, M.errorN, M.dt});
            if (hostI_[2 * b + 1]) { M.numIter = 2 * tgt; restartOrFail(b); continue; }
            M.numIter++;
        }
    } while (false);
    // a step that ended: advanceTimeLevel and the next step size
    // (suggestTimeStepSize); a member that failed for good stops
    for (int b = 0; b < B; ++b) {
        Member& M = *members_[b];
        if (state[b] == kFailedStep) { M.phase = 0; continue; }
        if (state[b] != kOk) continue;
        OM_HIP(hipMemcpyAsync(M.pv1.p, M.pv0.p, 3 * M.N * sizeof(double), hipMemcpyDeviceToDevice, stream_));
        OM_HIP(hipMemcpyAsync(M.mean1.p, M.mean0.p, M.N, hipMemcpyDeviceToDevice, stream_));
        M.dtHist.push_back(M.dt);
        M.stepsDone++;
        M.phase = 0;
        double nd;
        if (M.numIter > tgt) nd = std::max(prm_.min_dt, M.dt / (1.0 + double(M.numIter - tgt) / tgt));
        else nd = std::max(prm_.min_dt, M.dt * (1.0 + double(tgt - M.numIter) / tgt / 1.2));
        M.dt = std::min(prm_.max_dt, nd);
    }
    OM_HIP(hipStreamSynchronize(stream_));
}

bool Ensemble::allDone() const
{
    for (auto& m : members_) if (m->stepsDone < m->stepsTotal && !m->error) return false;
    return true;
}

template void Ensemble::bicgstab<double>(const std::vector<int>&, const std::vector<double>&, double, int,
                                         std::vector<int>&, std::vector<int>&);
template void Ensemble::bicgstab<float>(const std::vector<int>&, const std::vector<double>&, double, int,
                                        std::vector<int>&, std::vector<int>&);

// ---------------------------------------------------------------------------
Server::Server(const omodels_params& prm, double budgetBytes, int maxBatch)
    : prm_(prm), budget_(budgetBytes), maxBatch_(std::max(1, std::min(maxBatch, Ensemble::kMaxMembers))), ens_(prm)
{}

int Server::submit(const omodels_blackoil_problem* p, const double* pv, const uint8_t* mean, int nSteps, double dt0,
                   double arrival)
{
    Request r;
    r.id = (int)reqs_.size();
    r.prob = *p;
    const size_t N = (size_t)p->nx * p->ny * p->nz;
    r.perm.assign(p->perm, p->perm + 3 * N);
    r.poro.assign(p->poro, p->poro + N);
    r.oil.assign(p->oil_tab, p->oil_tab + 4 * p->n_oil);
    r.gas.assign(p->gas_tab, p->gas_tab + 3 * p->n_gas);
    r.srcCell.assign(p->src_cell, p->src_cell + p->n_src);
    r.srcType.assign(p->src_type, p->src_type + p->n_src);
    r.srcA.assign(p->src_a, p->src_a + p->n_src);
    r.srcB.assign(p->src_b, p->src_b + p->n_src);
    r.pv.assign(pv, pv + 3 * N);
    r.mean.assign(mean, mean + N);
    r.nSteps = nSteps;
    r.dt0 = dt0;
    r.arrival = arrival;
    r.bytes = Ensemble::memoryEstimate(p, prm_);
    if (r.bytes > budget_) throw std::invalid_argument("request does not fit the HBM budget");
    reqs_.push_back(std::move(r));
    return reqs_.back().id;
}

omodels_blackoil_problem Server::Request::view() const
{
    omodels_blackoil_problem p = prob;
    p.perm = perm.data(); p.poro = poro.data(); p.oil_tab = oil.data(); p.gas_tab = gas.data();
    p.src_cell = srcCell.data(); p.src_type = srcType.data(); p.src_a = srcA.data(); p.src_b = srcB.data();
    return p;
}

void Server::run()
{
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
    // the requests not run yet (a later call serves those submitted since)
    std::vector<int> order;
    for (size_t k = 0; k < reqs_.size(); ++k)
        if (reqs_[k].admit < 0.0) order.push_back((int)k);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs_[a].arrival < reqs_[b].arrival; });
    size_t next = 0;
    double used = 0.0;
    std::vector<std::pair<int, int>> running;   // (request, member id)
    int retired = 0;
    while (next < order.size() || !running.empty()) {
        // admission: arrived requests in FIFO order while memory and slots allow
        const double t = now();
        while (next < order.size() && reqs_[order[next]].arrival <= t && (int)running.size() < maxBatch_ &&
               used + reqs_[order[next]].bytes <= budget_) {
            Request& r = reqs_[order[next]];
            const omodels_blackoil_problem p = r.view();
            r.admit = now();
            const int mid = ens_.add(&p, r.pv.data(), r.mean.data(), r.nSteps, r.dt0);
            used += r.bytes;
            running.push_back({r.id, mid});
            r.pv.clear(); r.pv.shrink_to_fit();
            ++next;
        }
        maxRunning_ = std::max<int>(maxRunning_, (int)running.size());
        if (running.empty()) {
            const double wait = reqs_[order[next]].arrival - now();
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.05)));
            continue;
        }
        ens_.round();
        steps_++;
        // retire finished members
        for (size_t k = 0; k < running.size();) {
            Ensemble::Member* m = ens_.find(running[k].second);
            if (m->stepsDone >= m->stepsTotal || m->error) {
                Request& r = reqs_[running[k].first];
                r.finish = now();
                r.status = m->error ? m->error : 0;
                r.stats = m->stats;
                r.stats.t_total = r.finish - r.admit;
                r.result.resize(3 * m->N);
                r.resultMean.resize(m->N);
                ens_.download(m->id, r.result.data(), r.resultMean.data());
                r.log = m->log;
                r.dtHist = m->dtHist;
                used -= r.bytes;
                ens_.remove(m->id);
                running.erase(running.begin() + k);
                ++retired;
                if (prm_.verbose) {
                    printf("served request %d (%d of %zu) at %.1f s, latency %.2f s, status %d\n", r.id, retired,
                           reqs_.size(), r.finish, r.finish - r.arrival, r.status);
                    fflush(stdout);
                }
            } else ++k;
        }
    }
    wall_ = now();
}

}  // namespace omodels
