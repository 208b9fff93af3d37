// host_model.cpp -- Model / TpfaLinearizer, BiCGStab backend and the
// black-oil Newton method with its time loop, all driving the gfx950
// kernels on one HIP stream.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>

#include "omodels_host.h"

namespace omodels {

void hipCheck(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e) + " in " + what);
}

void sourceGroups(const int32_t* cell, int n, std::vector<int2>& groups, std::vector<int>& order)
{
    groups.clear();
    order.clear();
    std::vector<int> next(n, -1), firstOf;
    std::map<int32_t, std::pair<int, int>> cells;    // cell -> (first, last) source
    for (int k = 0; k < n; ++k) {
        auto it = cells.find(cell[k]);
        if (it == cells.end()) {
            cells.emplace(cell[k], std::make_pair(k, k));
            firstOf.push_back(k);
        } else {
            next[it->second.second] = k;
            it->second.second = k;
        }
    }
    for (int f : firstOf) {
        const int start = (int)order.size();
        for (int k = f; k >= 0; k = next[k]) order.push_back(k);
        groups.push_back(make_int2(start, (int)order.size() - start));
    }
}

// ---------------------------------------------------------------------------
Model::Model(const omodels_blackoil_problem* p, hipStream_t s)
{
    if (!p || p->nx < 1 || p->ny < 1 || p->nz < 1) throw std::invalid_argument("invalid grid");
    if (!p->perm || !p->poro || !p->oil_tab || !p->gas_tab) throw std::invalid_argument("missing problem arrays");
    initSystem(s);
    fluid = makeFluid(p);
    buildTopology(p, topo);
    N = topo.N;
    nnzb = topo.rowPtr[N];
    bs = 3;
    gx = p->nx; gy = p->ny; gz = p->nz;
    G.nx = p->nx; G.ny = p->ny; G.nz = p->nz;
    G.dx = p->dx; G.dy = p->dy; G.dz = p->dz; G.g = p->gravity;
    G.V = p->dx * p->dy * p->dz;
    G.area[0] = p->dy * p->dz; G.area[1] = p->dx * p->dz; G.area[2] = p->dx * p->dy;

    trans.upload(topo.trans.data(), topo.trans.size(), stream);
    transGrid = topo.trans;
    permHost.assign(p->perm, p->perm + 3 * (size_t)N);
    fluidDev.upload(&fluid, 1, stream);
    poro.upload(p->poro, N, stream);
    rowPtr.upload(topo.rowPtr.data(), topo.rowPtr.size(), stream);
    gridRowPtr.upload(topo.rowPtr.data(), topo.rowPtr.size(), stream);
    colIdx.upload(topo.colIdx.data(), topo.colIdx.size(), stream);
    diagOff.upload(topo.diagOff.data(), topo.diagOff.size(), stream);
    nbOff.upload(topo.nbOff.data(), topo.nbOff.size(), stream);
    levels.nLevels = (int)topo.levelPtr.size() - 1;
    levels.ptrHost.assign(topo.levelPtr.begin(), topo.levelPtr.end());
    levels.rows.upload(topo.levelRows.data(), topo.levelRows.size(), stream);
    nsrc = p->n_src;
    if (nsrc > 0) {
        for (int k = 0; k < nsrc; ++k)
            if (p->src_cell[k] < 0 || p->src_cell[k] >= N) throw std::invalid_argument("source cell out of range");
        srcCell.upload(p->src_cell, nsrc, stream);
        srcType.upload(p->src_type, nsrc, stream);
        srcA.upload(p->src_a, nsrc, stream);
        srcB.upload(p->src_b, nsrc, stream);
        std::vector<int2> groups;
        std::vector<int> order;
        sourceGroups(p->src_cell, nsrc, groups, order);
        nSrcGroups = (int)groups.size();
        srcGroups.upload(groups.data(), groups.size(), stream);
        srcOrder.upload(order.data(), order.size(), stream);
        sync();
    }
    pv0.alloc(3 * N); pv1.alloc(3 * N); st1.alloc(3 * N); res.alloc(3 * N);
    mean0.alloc(N); mean1.alloc(N); wasSw.alloc(N);
    OM_HIP(hipMemsetAsync(wasSw.p, 0, N, stream));
    OM_HIP(hipMemsetAsync(st1.p, 0, 3 * N * sizeof(double), stream));
    counters.alloc(4);
    ensureJacobian();
    sync();
}

void LinearSystem::initSystem(hipStream_t s)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("no HIP device available");
    if (s) stream = s;
    else { OM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)); ownStream = true; }
    partial.alloc(2 * kRedBlocks);
    dots.alloc(1);
    OM_HIP(hipMemsetAsync(dots.p, 0, sizeof(DotScratch), stream));
    scal.alloc(8);
    OM_HIP(hipHostMalloc((void**)&hostScal, 8 * sizeof(double), hipHostMallocDefault));
}

LinearSystem::~LinearSystem()
{
    if (stream) (void)hipStreamSynchronize(stream);
    if (hostScal) (void)hipHostFree(hostScal);
    if (ownStream && stream) (void)hipStreamDestroy(stream);
}

void LinearSystem::sync() { OM_HIP(hipStreamSynchronize(stream)); }

void Model::ensureJacobian()
{
    if (!jac.p) jac.alloc(9 * (size_t)nnzb);
}

// TpfaLinearizer::linearizeDomain (tpfalinearizer.hh:189-245) at solution(0)
void Model::linearizeDomain(double dt, int iterationIndex)
{
    ensureJacobian();
    // resetSystem_: with auxiliary rows the blocks the cell loop does not
    // write (the auxiliary entries) start from zero
    if (nAux > 0) {
        OM_HIP(hipMemsetAsync(jac.p, 0, 9 * (size_t)nnzb * sizeof(double), stream));
        OM_HIP(hipMemsetAsync(res.p, 0, 3 * (size_t)rows() * sizeof(double), stream));
    }
    if (iterationIndex == 0)
        launchBoStorage(fluid, N, pv0.p, mean0.p, poro.p, st1.p, stream, &geo.rock);
    const double storefac = G.V / dt;
    geo.dt = dt;
    timer.begin(kClsAssemble, stream);
    launchBoAssemble(fluidDev.p, G, pv0.p, mean0.p, poro.p, trans.p, st1.p, storefac, res.p, jac.p, diagOff.p,
                     nbOff.p, stream, &diffusion, &dispersion, &geo);
    timer.end(stream);
    // the velocities just written are the next linearization's input
    if (dispersion.disp) std::swap(dispersion.nvIn, dispersion.nvOut);
    // cell-wise sources (:827-841), or the same terms as sparse blocks added
    // after the cell loop (:842-846); both add each cell's term last
    if (!separateSparseSources)
        launchBoSources(fluid, G, pv0.p, mean0.p, nSrcGroups, srcGroups.p, srcOrder.p, srcCell.p, srcType.p, srcA.p,
                        srcB.p, res.p, jac.p, diagOff.p, stream, geo.vol);
    else
        addBuiltinSparseSources();
    // boundary terms after the cell loop (:849-865)
    launchBoBoundary(fluidDev.p, G, pv0.p, mean0.p, bc, geo.depth, res.p, jac.p, diagOff.p, stream);
    OM_HIP(hipGetLastError());
}

// linearizeDomain(domain) (tpfalinearizer.hh:226-245) on a part of the grid:
// resetSystem_(domain) clears the domain's rows (:370-379), then linearize_
// evaluates the domain's cells only (:711-842): their residual rows and
// diagonal blocks, and the column blocks A[J][I] added into the neighbours'
// rows (which are not cleared when the neighbour lies outside); the storage
// cache is not refreshed on a sub-domain (:793-803).  The separate sparse
// source terms and the boundary terms are added for all cells, as the
// reference's loops after the cell loop do (:843-865).  All cells: the full
// linearization.
void Model::linearizeSubdomain(double dt, int iterationIndex, int n, const int* cells)
{
    if (n <= 0 || !cells) throw std::invalid_argument("sub-domain: at least one cell");
    if (dispersion.disp || diffusion.diff) throw std::invalid_argument("sub-domain: not with diffusion or dispersion");
    std::vector<uint8_t> mask(N, 0);
    for (int k = 0; k < n; ++k) {
        if (cells[k] < 0 || cells[k] >= N) throw std::invalid_argument("sub-domain: cell out of range");
        if (mask[cells[k]]) throw std::invalid_argument("sub-domain: a cell given twice");
        mask[cells[k]] = 1;
    }
    if (n == rows()) { linearizeDomain(dt, iterationIndex); return; }
    ensureJacobian();
    sync();
    domMask.upload(mask.data(), mask.size(), stream);
    domCells.upload(cells, n, stream);
    launchClearRows(n, domCells.p, rowPtr.p, jac.p, res.p, stream);
    BoGeo g = geo;
    if (!g.depth) {
        if (!ownDepth.p) {
            const long nxy = (long)G.nx * G.ny;
            std::vector<double> z(N), v(N, G.V), th(3 * (size_t)N, 0.0);
            for (long I = 0; I < N; ++I) z[I] = G.dz * ((double)(I / nxy) + 0.5);
            ownDepth.upload(z.data(), z.size(), stream);
            ownVol.upload(v.data(), v.size(), stream);
            ownThp.upload(th.data(), th.size(), stream);
        }
        g.depth = ownDepth.p; g.vol = ownVol.p; g.thp = ownThp.p;
    }
    g.dt = dt;
    g.domain = domMask.p;
    timer.begin(kClsAssemble, stream);
    launchBoAssemble(fluidDev.p, G, pv0.p, mean0.p, poro.p, trans.p, st1.p, G.V / dt, res.p, jac.p, diagOff.p,
                     nbOff.p, stream, nullptr, nullptr, &g);
    timer.end(stream);
    if (!separateSparseSources)
        launchBoSources(fluid, G, pv0.p, mean0.p, nSrcGroups, srcGroups.p, srcOrder.p, srcCell.p, srcType.p, srcA.p,
                        srcB.p, res.p, jac.p, diagOff.p, stream, geo.vol, domMask.p);
    else
        addBuiltinSparseSources();
    launchBoBoundary(fluidDev.p, G, pv0.p, mean0.p, bc, geo.depth, res.p, jac.p, diagOff.p, stream);
    (void)iterationIndex;
    OM_HIP(hipGetLastError());
    sync();
}

void Model::setBoundary(int n, const int* cell, const int* face, const int* type, const double* rate,
                        const double* state, const uint8_t* meaning, const double* btrans, const double* zface)
{
    struct Face {
        int cell, face, type;
        double rate[3], state[3], trans, zf;
        uint8_t meaning;
    };
    std::vector<Face> fs;
    if (n < 0) throw std::invalid_argument("boundary: n >= 0");
    if (n > 0 && (!cell || !face || !type)) throw std::invalid_argument("boundary: cells, faces and types");
    const long nxy = (long)G.nx * G.ny;
    const double h[3] = {G.dx, G.dy, G.dz};
    for (int b = 0; b < n; ++b) {
        Face f{};
        f.cell = cell[b]; f.face = face[b]; f.type = type[b];
        if (f.cell < 0 || f.cell >= N || f.face < 0 || f.face > 5)
            throw std::invalid_argument("boundary: cell or face out of range");
        const long I = f.cell;
        const long ijk[3] = {I % G.nx, (I / G.nx) % G.ny, I / nxy};
        const int ext[3] = {G.nx, G.ny, G.nz};
        const int ax = f.face / 2;
        const bool lower = (f.face & 1) == 0;
        if (lower ? ijk[ax] > 0 : ijk[ax] + 1 < ext[ax]) throw std::invalid_argument("boundary: the face has a neighbour");
        if (f.type < 0 || f.type > 3) throw std::invalid_argument("boundary: type must be 0..3");
        if (f.type == 1) {
            if (!rate) throw std::invalid_argument("boundary: rate faces need rates");
            for (int e = 0; e < 3; ++e) {
                f.rate[e] = rate[3 * b + e];
                if (!std::isfinite(f.rate[e])) throw std::invalid_argument("boundary: rates must be finite");
            }
        }
        if (f.type >= 2) {
            if (!state || !meaning) throw std::invalid_argument("boundary: free faces need a state");
            for (int e = 0; e < 3; ++e) {
                f.state[e] = state[3 * b + e];
                if (!std::isfinite(f.state[e])) throw std::invalid_argument("boundary: states must be finite");
            }
            f.meaning = meaning[b];
            if (f.meaning != kMeanSg && f.meaning != kMeanRs) throw std::invalid_argument("boundary: meaning");
        }
        if (btrans) {
            f.trans = btrans[b];
            if (!(f.trans >= 0.0) || !std::isfinite(f.trans))
                throw std::invalid_argument("boundary: transmissibilities must be finite and >= 0");
        } else {
            // the cell's half of a face's transmissibility (transfluxmodule.hh:457-479)
            const double A = G.area[ax], ndot = 0.5 * h[ax], dist2 = ndot * ndot;
            f.trans = A * permHost[(size_t)ax * N + I] * ndot / dist2;
        }
        if (zface) {
            f.zf = zface[b];
            if (!std::isfinite(f.zf)) throw std::invalid_argument("boundary: face depths must be finite");
        } else {
            const double zc = depthHost.empty() ? G.dz * ((double)ijk[2] + 0.5) : depthHost[I];
            f.zf = ax < 2 ? zc : (f.face == 4 ? zc - 0.5 * G.dz : zc + 0.5 * G.dz);
        }
        fs.push_back(f);
    }
    std::stable_sort(fs.begin(), fs.end(),
                     [](const Face& a, const Face& b) { return a.cell != b.cell ? a.cell < b.cell : a.face < b.face; });
    for (size_t k = 1; k < fs.size(); ++k)
        if (fs[k].cell == fs[k - 1].cell && fs[k].face == fs[k - 1].face)
            throw std::invalid_argument("boundary: a face given twice");
    std::vector<int2> groups;
    std::vector<int> c(fs.size()), fa(fs.size()), ty(fs.size());
    std::vector<double> ra(3 * fs.size()), st(3 * fs.size()), tr(fs.size()), zf(fs.size());
    std::vector<uint8_t> me(fs.size());
    for (size_t k = 0; k < fs.size(); ++k) {
        if (k == 0 || fs[k].cell != fs[k - 1].cell) groups.push_back(make_int2((int)k, 0));
        groups.back().y++;
        c[k] = fs[k].cell; fa[k] = fs[k].face; ty[k] = fs[k].type;
        for (int e = 0; e < 3; ++e) { ra[3 * k + e] = fs[k].rate[e]; st[3 * k + e] = fs[k].state[e]; }
        tr[k] = fs[k].trans; zf[k] = fs[k].zf; me[k] = fs[k].meaning;
    }
    sync();
    bc = BoBc{};
    bcCellHost = c;
    bcFaceHost = fa;
    bcGroups.release(); bcCell.release(); bcFace.release(); bcType.release(); bcRate.release(); bcState.release();
    bcMean.release(); bcTrans.release(); bcZ.release();
    if (fs.empty()) return;
    bcGroups.upload(groups.data(), groups.size(), stream);
    bcCell.upload(c.data(), c.size(), stream);
    bcFace.upload(fa.data(), fa.size(), stream);
    bcType.upload(ty.data(), ty.size(), stream);
    bcRate.upload(ra.data(), ra.size(), stream);
    bcState.upload(st.data(), st.size(), stream);
    bcMean.upload(me.data(), me.size(), stream);
    bcTrans.upload(tr.data(), tr.size(), stream);
    bcZ.upload(zf.data(), zf.size(), stream);
    sync();
    bc = BoBc{(int)groups.size(), bcGroups.p, bcCell.p, bcFace.p, bcType.p, bcRate.p, bcState.p, bcMean.p,
              bcTrans.p, bcZ.p};
}

// the boundary faces' terms at solution(0) as FLOWS records them (the
// faces in cell and face order; type-0 faces zero)
int Model::boundaryFlows(double* hostFlows, int* hostCell, int* hostFace)
{
    const int n = (int)bcCellHost.size();
    if (n == 0) return 0;
    if (hostCell) std::copy(bcCellHost.begin(), bcCellHost.end(), hostCell);
    if (hostFace) std::copy(bcFaceHost.begin(), bcFaceHost.end(), hostFace);
    if (hostFlows) {
        DevBuf<double> f;
        f.alloc(3 * (size_t)n);
        OM_HIP(hipMemsetAsync(f.p, 0, 3 * (size_t)n * sizeof(double), stream));
        launchBoBoundary(fluidDev.p, G, pv0.p, mean0.p, bc, geo.depth, nullptr, nullptr, nullptr, stream, f.p);
        OM_HIP(hipGetLastError());
        f.download(hostFlows, 3 * (size_t)n, stream);
        sync();
    }
    return n;
}

void Model::computeSourceBlocks()
{
    srcRes.alloc(3 * (size_t)std::max(nsrc, 1));
    srcDiag.alloc(9 * (size_t)std::max(nsrc, 1));
    launchBoSourceBlocks(fluid, G, pv0.p, mean0.p, nsrc, srcCell.p, srcType.p, srcA.p, srcB.p, srcRes.p,
                         srcDiag.p, stream, geo.vol);
    OM_HIP(hipGetLastError());
}

void Model::addSparseSources(int n, const int* cells, const double* r, const double* d)
{
    ensureJacobian();
    launchAddSparseSources(n, cells, r, d, res.p, jac.p, diagOff.p, stream);
    OM_HIP(hipGetLastError());
}

void Model::computeFlows(double* flows, double* flores)
{
    timer.begin(kClsOther, stream);
    launchBoFlows(fluidDev.p, G, pv0.p, mean0.p, trans.p, gridRowPtr.p, flows, flores, stream, &diffusion, &dispersion,
                  &geo);
    timer.end(stream);
    OM_HIP(hipGetLastError());
}

namespace {
// harmonic combination of two half-face values; a face between two cells
// without porosity (diffusion) or without a dispersion coefficient has none:
// 0 rather than 0/0
double harmonicFace(double a, double b) { return a + b == 0.0 ? 0.0 : a * b / (a + b); }
void checkNonNegative(const double* v, long n, const char* what)
{
    for (long i = 0; i < n; ++i)
        if (!(v[i] >= 0.0) || !std::isfinite(v[i])) throw std::invalid_argument(what);
}
}  // namespace

void Model::updateDiscretization(const double* faceTrans, const double* depth, const double* volume,
                                 const double* thpres)
{
    if (faceTrans) checkNonNegative(faceTrans, 3 * N, "transmissibilities must be finite and >= 0");
    if (thpres) checkNonNegative(thpres, 3 * N, "threshold pressures must be finite and >= 0");
    if (depth)
        for (long I = 0; I < N; ++I)
            if (!std::isfinite(depth[I])) throw std::invalid_argument("depths must be finite");
    if (volume)
        for (long I = 0; I < N; ++I)
            if (!(volume[I] > 0.0) || !std::isfinite(volume[I]))
                throw std::invalid_argument("cell volumes must be finite and > 0");
    // faces without a neighbour keep 0, as the grid's own table has them
    const long nxy = (long)G.nx * G.ny;
    std::vector<double> t = transGrid;
    if (faceTrans)
        for (long I = 0; I < N; ++I) {
            const long ijk[3] = {I % G.nx, (I / G.nx) % G.ny, I / nxy};
            const int n[3] = {G.nx, G.ny, G.nz};
            for (int ax = 0; ax < 3; ++ax)
                if (ijk[ax] + 1 < n[ax]) t[(size_t)ax * N + I] = faceTrans[(size_t)ax * N + I];
        }
    sync();
    trans.upload(t.data(), t.size(), stream);
    topo.trans = t;
    geoGiven = depth || volume || thpres;
    depthHost.clear();
    geoDepth.release(); geoVol.release(); geoThp.release();
    if (geoGiven) {
        std::vector<double> z(N), v(N), th(3 * (size_t)N, 0.0);
        for (long I = 0; I < N; ++I) {
            z[I] = depth ? depth[I] : G.dz * ((double)(I / nxy) + 0.5);
            v[I] = volume ? volume[I] : G.V;
        }
        if (thpres) std::copy(thpres, thpres + 3 * N, th.begin());
        geoDepth.upload(z.data(), z.size(), stream);
        geoVol.upload(v.data(), v.size(), stream);
        geoThp.upload(th.data(), th.size(), stream);
        depthHost = z;
        sync();
    }
    refreshGeo();
}

void Model::refreshGeo()
{
    const bool rockOn = rockComp.p || rockTabs.p || dirMob.p;
    if (!geoGiven && !rockOn) {
        geoDepth.release(); geoVol.release(); geoThp.release();
        geo = BoGeo{};
        return;
    }
    if (!geoGiven && !geoDepth.p) {
        // the grid's own geometry for the kernels' extended form
        const long nxy = (long)G.nx * G.ny;
        std::vector<double> z(N), v(N, G.V), th(3 * (size_t)N, 0.0);
        for (long I = 0; I < N; ++I) z[I] = G.dz * ((double)(I / nxy) + 0.5);
        geoDepth.upload(z.data(), z.size(), stream);
        geoVol.upload(v.data(), v.size(), stream);
        geoThp.upload(th.data(), th.size(), stream);
        sync();
    }
    geo.depth = geoDepth.p; geo.vol = geoVol.p; geo.thp = geoThp.p;
    geo.rock = BoRock{rockComp.p, rockPref.p, rockTabs.p, rockTabs.p ? rockTabs.p + 1 : nullptr};
    geo.dirMob = dirMob.p;
}

void Model::setDirectionalMobility(const double* mult)
{
    if (mult) checkNonNegative(mult, 9 * N, "directional mobility multipliers must be finite and >= 0");
    sync();
    dirMob.release();
    if (mult) dirMob.upload(mult, 9 * (size_t)N, stream);
    sync();
    refreshGeo();
}

void Model::setRock(const double* comp, const double* pref, int nTab, const double* table)
{
    if (nTab < 0 || (nTab > 0 && !table)) throw std::invalid_argument("rock table: n >= 0 and a table when n > 0");
    if (comp) checkNonNegative(comp, N, "rock compressibilities must be finite and >= 0");
    if (pref)
        for (long I = 0; I < N; ++I)
            if (!std::isfinite(pref[I])) throw std::invalid_argument("rock reference pressures must be finite");
    Tab tabs[2];
    if (nTab > 0) {
        for (int k = 0; k < 2 * nTab; ++k)
            if (!(table[nTab + k] > 0.0) || !std::isfinite(table[nTab + k]))
                throw std::invalid_argument("rock multipliers must be finite and > 0");
        fillTab(tabs[0], table, table + nTab, nTab);
        fillTab(tabs[1], table, table + 2 * nTab, nTab);
    }
    sync();
    rockComp.release(); rockPref.release(); rockTabs.release();
    if (comp) rockComp.upload(comp, N, stream);
    if (comp && pref) rockPref.upload(pref, N, stream);
    if (nTab > 0) rockTabs.upload(tabs, 2, stream);
    sync();
    refreshGeo();
}

void Model::setDiffusion(const omodels_diffusion* d)
{
    if (!d) { diffusion = BoDiff{}; faceDiff.release(); return; }
    if (!(d->oil_in_oil >= 0.0) || !(d->gas_in_oil >= 0.0) || !std::isfinite(d->oil_in_oil) ||
        !std::isfinite(d->gas_in_oil))
        throw std::invalid_argument("diffusion coefficients must be finite and >= 0");
    std::vector<double> fd(3 * (size_t)N, 0.0);
    if (d->face_diffusivity) {
        checkNonNegative(d->face_diffusivity, 3 * N, "face diffusivities must be finite and >= 0");
        std::copy(d->face_diffusivity, d->face_diffusivity + 3 * N, fd.begin());
    } else {
        // harmonic mean of the half-face values A phi |n.d| / |d|^2
        std::vector<double> phi(N);
        poro.download(phi.data(), N, stream);
        sync();
        const double h[3] = {G.dx, G.dy, G.dz};
        const long stride[3] = {1, (long)G.nx, (long)G.nx * G.ny};
        for (long I = 0; I < N; ++I) {
            const long ijk[3] = {I % G.nx, (I / G.nx) % G.ny, I / ((long)G.nx * G.ny)};
            const int n[3] = {G.nx, G.ny, G.nz};
            for (int ax = 0; ax < 3; ++ax) {
                if (ijk[ax] + 1 >= n[ax]) continue;
                const long J = I + stride[ax];
                const double A = G.area[ax], ndot = 0.5 * h[ax], dist2 = ndot * ndot;
                const double D0 = A * phi[I] * ndot / dist2;
                const double D1 = A * phi[J] * ndot / dist2;
                fd[(size_t)ax * N + I] = harmonicFace(D0, D1);
            }
        }
    }
    faceDiff.upload(fd.data(), fd.size(), stream);
    const bool mole = d->mw_oil > 0.0 && d->mw_gas > 0.0;
    const double mO = mole ? d->mw_oil : 1.0, mG = mole ? d->mw_gas : 1.0;
    diffusion.toFrac = fluid.rho[kO] * mG / (fluid.rho[kG] * mO);
    diffusion.effOil = 0.5 * (d->oil_in_oil + d->oil_in_oil);
    diffusion.effGas = 0.5 * (d->gas_in_oil + d->gas_in_oil);
    diffusion.diff = faceDiff.p;
    sync();
}

void Model::setDispersion(const omodels_dispersion* d)
{
    if (!d) {
        dispersion = BoDisp{};
        faceDisp.release(); nvA.release(); nvB.release();
        return;
    }
    if (!d->face_dispersivity && !d->cell_dispersion)
        throw std::invalid_argument("dispersion needs face dispersivities or cell coefficients");
    std::vector<double> fd(3 * (size_t)N, 0.0);
    if (d->face_dispersivity) {
        checkNonNegative(d->face_dispersivity, 3 * N, "face dispersivities must be finite and >= 0");
        std::copy(d->face_dispersivity, d->face_dispersivity + 3 * N, fd.begin());
    } else {
        checkNonNegative(d->cell_dispersion, N, "dispersion coefficients must be finite and >= 0");
        // harmonic mean of the half-face values A E |n.d| / |d|^2
        const double* E = d->cell_dispersion;
        const double h[3] = {G.dx, G.dy, G.dz};
        const long stride[3] = {1, (long)G.nx, (long)G.nx * G.ny};
        const int n[3] = {G.nx, G.ny, G.nz};
        for (long I = 0; I < N; ++I) {
            const long ijk[3] = {I % G.nx, (I / G.nx) % G.ny, I / ((long)G.nx * G.ny)};
            for (int ax = 0; ax < 3; ++ax) {
                if (ijk[ax] + 1 >= n[ax]) continue;
                const long J = I + stride[ax];
                const double A = G.area[ax], ndot = 0.5 * h[ax], dist2 = ndot * ndot;
                const double E0 = A * E[I] * ndot / dist2;
                const double E1 = A * E[J] * ndot / dist2;
                fd[(size_t)ax * N + I] = harmonicFace(E0, E1);
            }
        }
    }
    faceDisp.upload(fd.data(), fd.size(), stream);
    nvA.alloc(N); nvB.alloc(N);
    OM_HIP(hipMemsetAsync(nvA.p, 0, N * sizeof(double), stream));
    OM_HIP(hipMemsetAsync(nvB.p, 0, N * sizeof(double), stream));
    dispersion.toMassFrac = fluid.rho[kO] / fluid.rho[kG];
    dispersion.disp = faceDisp.p;
    dispersion.nvIn = nvA.p;
    dispersion.nvOut = nvB.p;
    sync();
}

void Model::addBuiltinSparseSources()
{
    if (nsrc <= 0) return;
    computeSourceBlocks();
    addSparseSources(nsrc, srcCell.p, srcRes.p, srcDiag.p);
}

double LinearSystem::maxAbsResidual()
{
    launchMaxAbs<double>((long)bs * N, res.p, partial.p, scal.p, stream);
    OM_HIP(hipMemcpyAsync(hostScal, scal.p, sizeof(double), hipMemcpyDeviceToHost, stream));
    sync();
    return hostScal[0];
}

void Model::advanceTimeLevel()
{
    OM_HIP(hipMemcpyAsync(pv1.p, pv0.p, 3 * N * sizeof(double), hipMemcpyDeviceToDevice, stream));
    OM_HIP(hipMemcpyAsync(mean1.p, mean0.p, N, hipMemcpyDeviceToDevice, stream));
    if (nAux > 0)
        OM_HIP(hipMemcpyAsync(aux1.p, aux0.p, 3 * nAux * sizeof(double), hipMemcpyDeviceToDevice, stream));
}

void Model::restoreTimeLevel()
{
    OM_HIP(hipMemcpyAsync(pv0.p, pv1.p, 3 * N * sizeof(double), hipMemcpyDeviceToDevice, stream));
    OM_HIP(hipMemcpyAsync(mean0.p, mean1.p, N, hipMemcpyDeviceToDevice, stream));
    if (nAux > 0)
        OM_HIP(hipMemcpyAsync(aux0.p, aux1.p, 3 * nAux * sizeof(double), hipMemcpyDeviceToDevice, stream));
}

void buildLevelSchedule(long R, const std::vector<int>& rp, const std::vector<int>& ci, LevelSchedule& L,
                        hipStream_t s)
{
    // lower-numbered rows coupled through an entry above the diagonal
    std::vector<std::vector<int>> upFrom(R);
    for (long r = 0; r < R; ++r)
        for (int k = rp[r]; k < rp[r + 1]; ++k)
            if (ci[k] > r) upFrom[ci[k]].push_back((int)r);
    std::vector<int> lvl(R, 0);
    int nL = 0;
    for (long i = 0; i < R; ++i) {
        int l = 0;
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] < i) l = std::max(l, lvl[ci[k]] + 1);
        for (int j : upFrom[i]) l = std::max(l, lvl[j] + 1);
        lvl[i] = l;
        nL = std::max(nL, l + 1);
    }
    L.nLevels = nL;
    L.ptrHost.assign(nL + 1, 0);
    for (long i = 0; i < R; ++i) L.ptrHost[lvl[i] + 1]++;
    for (int l = 0; l < nL; ++l) L.ptrHost[l + 1] += L.ptrHost[l];
    std::vector<int> order(R), fill(L.ptrHost.begin(), L.ptrHost.end() - 1);
    for (long i = 0; i < R; ++i) order[fill[lvl[i]]++] = (int)i;
    L.rows.upload(order.data(), order.size(), s);
}

long Model::blockOffset(long row, long col) const
{
    const std::vector<int>& rp = nAux > 0 ? xRowPtr : topo.rowPtr;
    const std::vector<int>& ci = nAux > 0 ? xColIdx : topo.colIdx;
    if (row < 0 || row >= rows() || col < 0 || col >= rows()) return -1;
    const int* b = ci.data() + rp[row];
    const int* e = ci.data() + rp[row + 1];
    const int* it = std::lower_bound(b, e, (int)col);
    return (it == e || *it != col) ? -1 : (long)(it - ci.data());
}

// createMatrix_ with the auxiliary module's addNeighbors (tpfalinearizer.hh:
// 503-507): the module's DOFs follow the cells; every row's columns are the
// sorted union of its grid neighbours and the module's entries.
void Model::setAuxiliary(int n, int nLinks, const int* row, const int* col)
{
    if (n < 0 || nLinks < 0) throw std::invalid_argument("auxiliary: counts must be >= 0");
    if (n > 0 && (nLinks == 0 || !row || !col)) throw std::invalid_argument("auxiliary: the module's pattern entries");
    const long R = N + n;
    std::vector<int> rp(R + 1, 0), ci, dg(R), nb(6 * (size_t)N, -1);
    if (n > 0) {
        std::vector<std::vector<int>> extra(R);
        for (int k = 0; k < nLinks; ++k) {
            if (row[k] < 0 || row[k] >= R || col[k] < 0 || col[k] >= R)
                throw std::invalid_argument("auxiliary: entry out of range");
            extra[row[k]].push_back(col[k]);
        }
        std::vector<int> cols;
        for (long r = 0; r < R; ++r) {
            cols.clear();
            if (r < N) cols.assign(topo.colIdx.begin() + topo.rowPtr[r], topo.colIdx.begin() + topo.rowPtr[r + 1]);
            cols.insert(cols.end(), extra[r].begin(), extra[r].end());
            std::sort(cols.begin(), cols.end());
            cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
            if (r >= N && !std::binary_search(cols.begin(), cols.end(), (int)r))
                throw std::invalid_argument("auxiliary: a row without its diagonal block");
            ci.insert(ci.end(), cols.begin(), cols.end());
            rp[r + 1] = (int)ci.size();
        }
        if (ci.size() >= (size_t)INT32_MAX / 9) throw std::invalid_argument("auxiliary: pattern too large");
        auto at = [&](long r, long c) {
            const int* it = std::lower_bound(ci.data() + rp[r], ci.data() + rp[r + 1], (int)c);
            return (int)(it - ci.data());
        };
        for (long r = 0; r < R; ++r) dg[r] = at(r, r);
        const long nxy = (long)G.nx * G.ny;
        for (int d = 0; d < 6; ++d)
            for (long I = 0; I < N; ++I) {
                if (topo.nbOff[(long)d * N + I] < 0) continue;
                const long J = d == 0 ? I - 1 : d == 1 ? I + 1 : d == 2 ? I - G.nx : d == 3 ? I + G.nx
                             : d == 4 ? I - nxy : I + nxy;
                nb[(long)d * N + I] = at(J, I);    // the block A[J][I]
            }
    }
    sync();
    nAux = n;
    if (n > 0) {
        xRowPtr = std::move(rp); xColIdx = std::move(ci); xDiagOff = std::move(dg); xNbOff = std::move(nb);
        rowPtr.upload(xRowPtr.data(), xRowPtr.size(), stream);
        colIdx.upload(xColIdx.data(), xColIdx.size(), stream);
        diagOff.upload(xDiagOff.data(), xDiagOff.size(), stream);
        nbOff.upload(xNbOff.data(), xNbOff.size(), stream);
        nnzb = xRowPtr[R];
        buildLevelSchedule(R, xRowPtr, xColIdx, levels, stream);
        aux0.alloc(3 * (size_t)n); aux1.alloc(3 * (size_t)n);
        OM_HIP(hipMemsetAsync(aux0.p, 0, 3 * (size_t)n * sizeof(double), stream));
        OM_HIP(hipMemsetAsync(aux1.p, 0, 3 * (size_t)n * sizeof(double), stream));
    } else {
        xRowPtr.clear(); xColIdx.clear(); xDiagOff.clear(); xNbOff.clear();
        rowPtr.upload(topo.rowPtr.data(), topo.rowPtr.size(), stream);
        colIdx.upload(topo.colIdx.data(), topo.colIdx.size(), stream);
        diagOff.upload(topo.diagOff.data(), topo.diagOff.size(), stream);
        nbOff.upload(topo.nbOff.data(), topo.nbOff.size(), stream);
        nnzb = topo.rowPtr[N];
        levels.nLevels = (int)topo.levelPtr.size() - 1;
        levels.ptrHost.assign(topo.levelPtr.begin(), topo.levelPtr.end());
        levels.rows.upload(topo.levelRows.data(), topo.levelRows.size(), stream);
        aux0.release(); aux1.release();
    }
    jac.release();
    ensureJacobian();
    res.alloc(3 * (size_t)rows());
    OM_HIP(hipMemsetAsync(res.p, 0, 3 * (size_t)rows() * sizeof(double), stream));
    OM_HIP(hipMemsetAsync(jac.p, 0, 9 * (size_t)nnzb * sizeof(double), stream));
    sync();
}

void Model::addAuxiliary(int nb, const int* blockOff, const double* blocks, int nr, const int* resRow,
                         const double* resVals)
{
    ensureJacobian();
    launchAddAuxiliary(nb, blockOff, blocks, nr, resRow, resVals, jac.p, res.p, stream);
    OM_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
BiCGStabBackend::BiCGStabBackend(LinearSystem& m, const omodels_params& prm)
    : m_(m), builtAux_(m.nAux), builtNnzb_(m.nnzb), prm_(prm)
{
    const long n = (long)m.bs * m.rows();
    if (prm_.precond < 0 || prm_.precond > 2) throw std::invalid_argument("precond must be 0, 1 or 2");
    if (prm_.precond == 1 && m.bs != 3) throw std::invalid_argument("block Jacobi is built for 3x3 blocks");
    b64_.alloc(n); x64_.alloc(n);
    ks_.alloc(1);
    fail_.alloc(1);
    OM_HIP(hipHostMalloc((void**)&ksHost_, 2 * sizeof(KrylovScalars), hipHostMallocDefault));
    for (auto& e : ksEv_) OM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (prm_.precision < 0 || prm_.precision > 2) throw std::invalid_argument("precision must be 0, 1 or 2");
    if (prm_.precision == 2 && prm_.precond != 0) throw std::invalid_argument("the FP32 solve uses block ILU(0)");
    if (prm_.ilu_schedule < 0 || prm_.ilu_schedule > 3) throw std::invalid_argument("ilu_schedule must be 0..3");
    if (prm_.ilu_order < 0 || prm_.ilu_order > 1) throw std::invalid_argument("ilu_order must be 0 or 1");
    // the tile and two-colour sweeps are built for grid systems; auxiliary
    // rows take the level-scheduled sweeps
    const bool tiles = prm_.ilu_schedule != 1 && m.nAux == 0;
    if (prm_.precond == 0 && prm_.ilu_order == 1) {
        if (m.bs != 3 || prm_.precision == 2 || m.nAux > 0)
            throw std::invalid_argument("the two-colour ILU(0) is built for 3x3 blocks and FP64 pivots");
        colored_ = std::make_unique<ColoredIlu>(m.gx, m.gy, m.gz, m.stream);
    } else if (prm_.precond == 0 && tiles && m.bs == 3 && prm_.precision != 2) {
        TunedIlu t;
        if (prm_.autotune && prm_.ilu_schedule == 0 && lookupTunedIlu(m.gx, m.gy, m.gz, prm_.precision, t)) {
            tuned_ = true;
            if (!t.level) tiled_ = std::make_unique<TiledIlu>(t.shape, m.stream, t.sweep);
        } else {
            const int mode = prm_.ilu_schedule == 2 ? 1 : (prm_.ilu_schedule == 3 ? 2 : 0);
            tiled_ = std::make_unique<TiledIlu>(makeTileGeom(m.gx, m.gy, m.gz), m.stream, mode);
        }
    }
    if (prm_.precond == 1) {
        if (prm_.prec_order < 0) throw std::invalid_argument("PreconditionerOrder must be >= 0");
        jacTmp_.alloc(n * (prm_.precision == 0 ? sizeof(double) : sizeof(float)));
    }
    if (prm_.precond == 2) {
        if (!structured()) throw std::invalid_argument("the AMG is built for 3x3 systems on a 7-point grid");
        if (prm_.precision == 2) throw std::invalid_argument("the AMG runs the FP64 solve or the FP32 inner solve");
        AmgParams ap;
        ap.coarsenTarget = prm_.amg_coarsen_target;
        ap.pre = prm_.amg_pre_smooth;
        ap.post = prm_.amg_post_smooth;
        ap.damp = prm_.amg_prolong_damp;
        for (int q = 0; q < 3; ++q) ap.agg[q] = prm_.amg_agg[q];
        amg_ = std::make_unique<AmgHierarchy>(m.gx, m.gy, m.gz, ap, prm_.precision == 1, m.stream);
    }
    if (prm_.precision == 0) {
        for (auto& w : wd_) w.alloc(n);
        r64_.alloc(n);
    } else {
        for (auto& w : wf_) w.alloc(n);
        r64_.alloc(n);
        bf_.alloc(n); xf_.alloc(n);
    }
}

BiCGStabBackend::~BiCGStabBackend()
{
    for (auto e : ksEv_)
        if (e) (void)hipEventDestroy(e);
    if (ksHost_) (void)hipHostFree(ksHost_);
}

// prepare(): ParallelBaseBackend::prepare/setResidual/setMatrix
// (parallelbasebackend.hh:182-250) + preparePreconditioner_ (:311-333):
// the solver-owned copy in the solver scalar and the preconditioner factors.
// Autotuner: time the factorization and applications of every tile shape
// (and of the level-scheduled kernels) on the current matrix, in two
// interleaved rounds, and keep the fastest by factor + 50 applications (a
// factorization serves about that many preconditioner applications in these
// solves), using the median of five applications per round and the better of
// the two rounds.  The default shape is kept unless another one is more than
// 5% faster.  Every candidate applies each row's operations in the same
// order, so the choice never changes results.
void BiCGStabBackend::tuneIlu()
{
    tuned_ = true;
    hipStream_t s = m_.stream;
    const int N = (int)m_.rows();
    const long nnzb = m_.nnzb, n = 3L * N;
    DevBuf<double> lu, dinv, v64;
    DevBuf<float> luf, df, vf;
    lu.alloc(9 * nnzb); dinv.alloc(9 * (size_t)N); v64.alloc(n);
    if (prm_.precision == 1) { luf.alloc(9 * nnzb); df.alloc(n); vf.alloc(n); launchCopyConvertD<float>(n, b64_.p, df.p, s); }
    constexpr int kApply = 5;
    hipEvent_t ev[kApply + 3];
    for (auto& e : ev) OM_HIP(hipEventCreate(&e));
    auto run = [&](TiledIlu* t) {
        OM_HIP(hipMemcpyAsync(lu.p, A64_, 9 * nnzb * sizeof(double), hipMemcpyDeviceToDevice, s));
        OM_HIP(hipMemsetAsync(fail_.p, 0, sizeof(int), s));
        OM_HIP(hipEventRecord(ev[0], s));
        if (t) t->factor(m_.rowPtr.p, m_.diagOff.p, lu.p, dinv.p, fail_.p, s);
        else {
            launchIlu0Factor<double>(N, 3, m_.rowPtr.p, m_.colIdx.p, m_.diagOff.p, lu.p, m_.levels, fail_.p, s);
            launchExtractInvDiag<double>(N, 3, m_.diagOff.p, lu.p, dinv.p, s);
        }
        OM_HIP(hipEventRecord(ev[1], s));
        if (prm_.precision == 1) launchCopyConvertD<float>(9 * nnzb, lu.p, luf.p, s);
        for (int r = 0; r <= kApply; ++r) {   // application 0 warms up
            if (prm_.precision == 1) {
                if (t) t->apply<float, double>(m_.rowPtr.p, m_.diagOff.p, luf.p, dinv.p, df.p, vf.p, s);
                else launchIlu0Apply<float, double>(N, 3, m_.rowPtr.p, m_.colIdx.p, m_.diagOff.p, luf.p, dinv.p,
                                                    df.p, vf.p, m_.levels, s);
            } else {
                if (t) t->apply<double, double>(m_.rowPtr.p, m_.diagOff.p, lu.p, dinv.p, b64_.p, v64.p, s);
                else launchIlu0Apply<double, double>(N, 3, m_.rowPtr.p, m_.colIdx.p, m_.diagOff.p, lu.p, dinv.p,
                                                      b64_.p, v64.p, m_.levels, s);
            }
            OM_HIP(hipEventRecord(ev[2 + r], s));
        }
        OM_HIP(hipEventSynchronize(ev[2 + kApply]));
        if (t) t->checkError(s);
        float fms = 0.f;
        OM_HIP(hipEventElapsedTime(&fms, ev[0], ev[1]));
        std::vector<double> a(kApply);
        for (int r = 0; r < kApply; ++r) {
            float ms = 0.f;
            OM_HIP(hipEventElapsedTime(&ms, ev[2 + r], ev[3 + r]));
            a[r] = ms;
        }
        std::sort(a.begin(), a.end());
        return (double)fms + 50.0 * a[kApply / 2];
    };
    // candidates: every tile shape in both sweep forms (streaming, batched), then
    // the level schedule; candidate 0 is the default shape in the form the
    // default rule picks for it
    const std::vector<TileGeom> shapes = tileShapeCandidates(m_.gx, m_.gy, m_.gz);
    struct Cand { int shape, sweep; };
    std::vector<Cand> cands;
    {
        TiledIlu probe(shapes[0], s);
        const int first = probe.streaming() ? 1 : 2;
        cands.push_back({0, first});
        cands.push_back({0, 3 - first});
    }
    for (size_t c = 1; c < shapes.size(); ++c) {
        cands.push_back({(int)c, 1});
        cands.push_back({(int)c, 2});
    }
    std::vector<double> cost(cands.size() + 1, 1e300);   // last entry: level schedule
    for (int round = 0; round < 2; ++round) {
        for (size_t c = 0; c < cands.size(); ++c) {
            TiledIlu t(shapes[cands[c].shape], s, cands[c].sweep);
            cost[c] = std::min(cost[c], run(&t));
        }
        cost.back() = std::min(cost.back(), run(nullptr));
    }
    for (auto& e : ev) OM_HIP(hipEventDestroy(e));
    size_t best = 0;
    for (size_t c = 1; c < cost.size(); ++c)
        if (cost[c] < cost[best]) best = c;
    if (cost[0] <= 1.05 * cost[best]) best = 0;
    if (prm_.verbose) {
        for (size_t c = 0; c < cands.size(); ++c)
            printf("autotune ILU tile %dx%dx%d %s: %.3f ms\n", shapes[cands[c].shape].BX, shapes[cands[c].shape].BY,
                   shapes[cands[c].shape].BZ, cands[c].sweep == 1 ? "streaming" : "batched", cost[c]);
        printf("autotune ILU level schedule: %.3f ms; chosen: %s\n", cost.back(),
               best == cands.size() ? "level" : "tiles");
    }
    TunedIlu choice;
    choice.level = best == cands.size();
    if (!choice.level) {
        choice.shape = shapes[cands[best].shape];
        choice.sweep = cands[best].sweep;
    }
    storeTunedIlu(m_.gx, m_.gy, m_.gz, prm_.precision, choice);
    if (choice.level) tiled_.reset();
    else tiled_ = std::make_unique<TiledIlu>(choice.shape, s, choice.sweep);
}

void BiCGStabBackend::prepare(const double* hostValues, const double* hostResidual)
{
    hipStream_t s = m_.stream;
    const int bs = m_.bs, bb = bs * bs;
    const long n = (long)bs * m_.rows(), nnzb = m_.nnzb;
    // a solver is sized for the system it was created on (auxiliary DOFs
    // change the number of rows)
    if ((long)b64_.n != n || m_.nAux != builtAux_ || m_.nnzb != builtNnzb_)
        throw std::logic_error("the system's rows changed since the solver was created: create it again");
    if (hostValues) { Acopy_.upload(hostValues, bb * nnzb, s); A64_ = Acopy_.p; }
    else A64_ = m_.jac.p;
    if (hostResidual) b64_.upload(hostResidual, n, s);
    else OM_HIP(hipMemcpyAsync(b64_.p, m_.res.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (prm_.autotune && prm_.ilu_schedule == 0 && tiled_ && !tuned_) tuneIlu();
    OM_HIP(hipMemsetAsync(fail_.p, 0, sizeof(int), s));
    const int N = (int)m_.rows();
    // the inner solve's operator (and the two-colour sweeps) read the matrix in
    // the structured 7-slot layout of a 3x3 7-point grid system: whole cache
    // lines, no column indices, the same bits as the BCRS product
    if (structured() && prm_.precision != 2) {
        const long nd = 7L * bb * N;
        if (prm_.precision == 1) {
            if (amg_) {
                // the AMG hierarchy is formed in FP64
                dia64_.alloc(nd);
                launchToDia<double>(m_.gx, m_.gy, m_.gz, m_.rowPtr.p, A64_, dia64_.p, s);
            } else {
                dia64_.release();   // an FP64 copy made for a test application would be stale
            }
            diaf_.alloc(nd);
            launchToDia<float>(m_.gx, m_.gy, m_.gz, m_.rowPtr.p, A64_, diaf_.p, s);
        } else {
            dia64_.alloc(nd);
            launchToDia<double>(m_.gx, m_.gy, m_.gz, m_.rowPtr.p, A64_, dia64_.p, s);
        }
    }
    if (prm_.precision == 2) {
        // the reference's LinearSolverScalar = float: matrix, factorization and
        // Krylov solve all in FP32 (the inverted FP32 pivots are kept as FP64
        // numbers, which represents them exactly)
        Af_.alloc(bb * nnzb);
        launchCopyConvertD<float>(bb * nnzb, A64_, Af_.p, s);
        luf_.alloc(bb * nnzb);
        launchCopyConvertD<float>(bb * nnzb, A64_, luf_.p, s);
        dinv64_.alloc(bb * (size_t)N);
        m_.timer.begin(kClsIluFactor, s);
        launchIlu0Factor<float>(N, bs, m_.rowPtr.p, m_.colIdx.p, m_.diagOff.p, luf_.p, m_.levels, fail_.p, s);
        launchExtractInvDiag<float>(N, bs, m_.diagOff.p, luf_.p, dinv64_.p, s);
        m_.timer.end(s);
    } else if (amg_) {
        // the AMG hierarchy of the structured FP64 matrix
        m_.timer.begin(kClsIluFactor, s);
        amg_->setup(dia64_.p, fail_.p, s);
        m_.timer.end(s);
    } else if (colored_) {
        // two-colour ILU(0): FP64 pivots only; the sweeps read the solver's
        // matrix in the structured layout (FP64, or its FP32 copy)
        dinv64_.alloc(bb * (size_t)N);
        m_.timer.begin(kClsIluFactor, s);
        if (prm_.precision == 1) colored_->factor(m_.rowPtr.p, A64_, dinv64_.p, fail_.p, s);
        else colored_->factor(nullptr, dia64_.p, dinv64_.p, fail_.p, s);
        m_.timer.end(s);
    } else if (prm_.precond == 0) {
        // ILU(0) factorization in FP64 (pivot blocks stay FP64 in both forms)
        lu64_.alloc(bb * nnzb);
        OM_HIP(hipMemcpyAsync(lu64_.p, A64_, bb * nnzb * sizeof(double), hipMemcpyDeviceToDevice, s));
        dinv64_.alloc(bb * (size_t)N);
        m_.timer.begin(kClsIluFactor, s);
        if (tiled_) {
            tiled_->factor(m_.rowPtr.p, m_.diagOff.p, lu64_.p, dinv64_.p, fail_.p, s);
        } else {
            launchIlu0Factor<double>(N, bs, m_.rowPtr.p, m_.colIdx.p, m_.diagOff.p, lu64_.p, m_.levels, fail_.p, s);
            launchExtractInvDiag<double>(N, bs, m_.diagOff.p, lu64_.p, dinv64_.p, s);
        }
        m_.timer.end(s);
        if (prm_.precision == 1) {
            luf_.alloc(bb * nnzb);
            launchCopyConvertD<float>(bb * nnzb, lu64_.p, luf_.p, s);
        }
    } else {
        dinv64_.alloc(bb * (size_t)N);
        launchJacobiInvert<double>(N, bs, m_.diagOff.p, A64_, dinv64_.p, fail_.p, s);
    }
    if (prm_.precision == 1) {
        Af_.alloc(bb * nnzb);
        launchCopyConvertD<float>(bb * nnzb, A64_, Af_.p, s);
    }
    OM_HIP(hipGetLastError());
    int f = 0;
    OM_HIP(hipMemcpyAsync(&f, fail_.p, sizeof(int), hipMemcpyDeviceToHost, s));
    m_.sync();
    if (tiled_) tiled_->checkError(s);
    prepared_ = (f == 0);
}

template <class S>
bool BiCGStabBackend::precApply(const S* lu, const void* dinv, const S* d, S* v, const int* run, const S* A, S* Av)
{
    hipStream_t s = m_.stream;
    const int N = (int)m_.rows();
    bool fused = false;
    if (amg_) {
        m_.timer.begin(kClsIluApply, s);
        amg_->apply<S>(d, v, run, s);
        m_.timer.end(s);
    } else if (colored_) {
        if (!diaOf<S>()) throw std::logic_error("two-colour ILU: no structured matrix of this precision");
        m_.timer.begin(kClsIluApply, s);
        colored_->apply<S>(nullptr, diaOf<S>(), (const double*)dinv, d, v, prm_.prec_relax, s, run);
        m_.timer.end(s);
    } else if (prm_.precond == 0) {
        m_.timer.begin(kClsIluApply, s);
        if (tiled_ && A && prm_.spmv_in_sweep)
            fused = tiled_->apply<S, double>(m_.rowPtr.p, m_.diagOff.p, lu, (const double*)dinv, d, v, s, run,
                                             m_.colIdx.p, A, Av);
        else if (tiled_)
            tiled_->apply<S, double>(m_.rowPtr.p, m_.diagOff.p, lu, (const double*)dinv, d, v, s, run);
        else
            launchIlu0Apply<S, double>(N, m_.bs, m_.rowPtr.p, m_.colIdx.p, m_.diagOff.p, lu, (const double*)dinv,
                                       d, v, m_.levels, s);
        m_.timer.end(s);
    } else {
        (void)lu;
        // block Jacobi (Dune::SeqJac(A, PreconditionerOrder, w)): that many
        // sweeps on v from its current content, FP64 inverted pivots
        const S* Am = sizeof(S) == 8 ? (const S*)A64_ : (const S*)Af_.p;
        S* tmp = (S*)jacTmp_.p;
        launchJacobiSweeps<S>(N, prm_.prec_order, m_.rowPtr.p, m_.colIdx.p, Am, (const double*)dinv, d, v, tmp,
                              prm_.prec_relax, run, s);
    }
    return fused;
}

void BiCGStabBackend::IterGraphs::reset()
{
    if (first) (void)hipGraphExecDestroy(first);
    if (iter) (void)hipGraphExecDestroy(iter);
    first = iter = nullptr;
    generation = -1;
}

// BiCGStabSolver::apply (bicgstabsolver.hh:120-330) with CombinedCriterion;
// one host synchronisation per iteration (the convergence test).  The device
// scalars carry the iteration's tag base (2 it): the reductions of the first
// half tag their failures with 2 it, those of the second half with 2 it + 1.
template <class S>
int BiCGStabBackend::bicgstab(const S* A, const S* lu, const void* dinv, const S* b, S* x, double tol,
                              double absTol, int maxIter, int* its)
{
    hipStream_t s = m_.stream;
    const long m = (long)m_.bs * m_.rows();
    const int N = (int)m_.rows();
    S* r; S* p; S* v; S* y; S* z;
    if constexpr (sizeof(S) == 8) { r = (S*)wd_[0].p; p = (S*)wd_[1].p; v = (S*)wd_[2].p; y = (S*)wd_[3].p; z = (S*)wd_[4].p; }
    else { r = (S*)wf_[0].p; p = (S*)wf_[1].p; v = (S*)wf_[2].p; y = (S*)wf_[3].p; z = (S*)wf_[4].p; }
    OM_HIP(hipMemsetAsync(x, 0, m * sizeof(S), s));
    OM_HIP(hipMemcpyAsync(r, b, m * sizeof(S), hipMemcpyDeviceToDevice, s));
    // setInitial
    launchMaxAbs<S>(m, r, m_.partial.p, m_.scal.p, s);
    OM_HIP(hipMemcpyAsync(m_.hostScal, m_.scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
    m_.sync();
    const double minS = (double)std::numeric_limits<S>::min() * 1e10;
    const double init = std::max(m_.hostScal[0], minS);
    *its = 0;
    auto converged = [&](double res) { return res / std::max(1e-20, init) <= tol || res <= absTol; };
    if (converged(init)) return 0;
    KrylovScalars k0{};
    k0.rho = 1; k0.alpha = 1; k0.omega = 1; k0.init_norm = init;
    k0.fail_code = 0; k0.fail_tag = -1; k0.stag_half = 0; k0.stag_full_tag = -1;
    k0.tag_base = 0;
    k0.tol = tol; k0.abs_tol = absTol; k0.max_error = prm_.lin_max_error;
    k0.run = 1; k0.stop_rc = -1; k0.stop_back = 0;
    const int* run = reinterpret_cast<const int*>(reinterpret_cast<const char*>(ks_.p) + offsetof(KrylovScalars, run));
    OM_HIP(hipMemcpyAsync(ks_.p, &k0, sizeof(k0), hipMemcpyHostToDevice, s));
    OM_HIP(hipMemsetAsync(v, 0, m * sizeof(S), s));
    OM_HIP(hipMemsetAsync(p, 0, m * sizeof(S), s));
    // y and z start as copies of x = 0 (bicgstabsolver.hh:187-191); block
    // Jacobi sweeps start from their content
    OM_HIP(hipMemsetAsync(y, 0, m * sizeof(S), s));
    OM_HIP(hipMemsetAsync(z, 0, m * sizeof(S), s));
    // the last workgroup of every scalar product clears the ticket and the
    // accumulators; this clears what a solve interrupted by an error may have
    // left behind
    OM_HIP(hipMemsetAsync(m_.dots.p, 0, sizeof(DotScratch), s));

    // first half: rho, p, y = M^-1 p, v = A y, alpha, x/r update, convergence test.
    // Every kernel of the solve returns at once after the test has ended it
    // (KrylovScalars::run), so the host queues iteration it+1 before it has
    // seen the test of iteration it, and the device never waits for the host.
    bool fusedSpmv = false;     // the operator runs inside the preconditioner's backward sweep
    auto firstHalf = [&]() {
        launchDotFinal<S>(m, b, r, m_.dots.p, ks_.p, 0, 0, s);
        launchUpdateP<S>(m, r, v, p, ks_.p, s);
        fusedSpmv = precApply<S>(lu, dinv, p, y, run, A, v);
        if (!fusedSpmv) {
            m_.timer.begin(kClsSpmv, s);
            spmvInner<S>(A, y, v, run);
            m_.timer.end(s);
        }
        launchDotFinal<S>(m, b, v, m_.dots.p, ks_.p, 1, 0, s);
        launchUpdateXR<S>(m, y, v, x, r, m_.partial.p, ks_.p, s);
    };
    // second half: z = M^-1 r, y = A z, omega, x/r update, next tag base
    auto secondHalf = [&]() {
        fusedSpmv = precApply<S>(lu, dinv, r, z, run, A, y);
        if (!fusedSpmv) {
            m_.timer.begin(kClsSpmv, s);
            spmvInner<S>(A, z, y, run);
            m_.timer.end(s);
        }
        launchDot2Final<S>(m, y, r, m_.dots.p, ks_.p, 1, s);
        launchUpdateXR2<S>(m, z, y, x, r, m_.partial.p, ks_.p, 1, 2, s);
    };

    IterGraphs* gr = nullptr;
    if (prm_.graph) {
        gr = &graphs_[sizeof(S) == 4 ? 1 : 0];
        const std::array<const void*, 12> key = {A, lu, dinv, b, x, r, p, v, y, z, m_.partial.p, ks_.p};
        // the tiled ILU's device descriptors of both applications must be
        // resident before a capture (an upload cannot be recorded); a cache
        // turnover between the two primes is followed by one more pass
        if (tiled_ && prm_.precond == 0) {
            for (int pass = 0; pass < 2; ++pass) {
                const long g0 = tiled_->generation();
                tiled_->primeApply<S, double>(m_.rowPtr.p, m_.diagOff.p, lu, (const double*)dinv, p, y, s, run);
                tiled_->primeApply<S, double>(m_.rowPtr.p, m_.diagOff.p, lu, (const double*)dinv, r, z, s, run);
                if (tiled_->generation() == g0) break;
            }
        }
        const long gen = tiled_ ? tiled_->generation() : 0;
        if (gr->key != key || gr->generation != gen || !gr->first) {
            gr->reset();
            m_.sync();
            const bool timed = m_.timer.enabled;
            m_.timer.enabled = false;   // event records stay out of the graphs
            try {
                gr->first = captureGraph(s, firstHalf);
                gr->iter = captureGraph(s, [&]() { secondHalf(); firstHalf(); });
            } catch (...) {
                m_.timer.enabled = timed;
                gr->reset();
                throw;
            }
            m_.timer.enabled = timed;
            gr->key = key;
            gr->generation = gen;
        }
    }

    // iteration it: the second half of it-1 and the first half of it, then the
    // scalars (with the device's verdict) to host copy it % 2
    auto enqueue = [&](int it) {
        if (gr) {
            OM_HIP(hipGraphLaunch(it == 0 ? gr->first : gr->iter, s));
        } else {
            if (it > 0) secondHalf();
            firstHalf();
        }
        OM_HIP(hipMemcpyAsync(ksHost_ + (it & 1), ks_.p, sizeof(KrylovScalars), hipMemcpyDeviceToHost, s));
        OM_HIP(hipEventRecord(ksEv_[it & 1], s));
    };
    if (maxIter > 0) enqueue(0);
    for (int it = 0; it < maxIter; ++it) {
        if (it + 1 < maxIter) enqueue(it + 1);
        OM_HIP(hipEventSynchronize(ksEv_[it & 1]));
        const KrylovScalars K = ksHost_[it & 1];
        if (K.run == 0) {
            *its = it - K.stop_back;
            if (it + 1 < maxIter && !gr) {      // the queued iteration it+1 was skipped: untime it
                if (prm_.precond == 0) m_.timer.dropLast(kClsIluApply, 2);
                if (!fusedSpmv) m_.timer.dropLast(kClsSpmv, 2);
            }
            return K.stop_rc;
        }
        *its = it;
    }
    if (maxIter > 0) secondHalf();
    OM_HIP(hipMemcpyAsync(ksHost_, ks_.p, sizeof(KrylovScalars), hipMemcpyDeviceToHost, s));
    m_.sync();
    *its = maxIter;
    if (ksHost_->fail_code) return 2;
    return 1;
}

template <class S>
void BiCGStabBackend::spmvInner(const S* A, const S* x, S* y, const int* run)
{
    if (diaOf<S>()) launchSpmvDia<S>(m_.gx, m_.gy, m_.gz, diaOf<S>(), x, y, m_.stream, run);
    else launchSpmv<S>((int)m_.rows(), m_.bs, m_.rowPtr.p, m_.colIdx.p, A, x, y, m_.stream, run);
}

bool BiCGStabBackend::structured() const
{
    return m_.bs == 3 && m_.nAux == 0 && m_.gx > 0 && (long)m_.gx * m_.gy * m_.gz == m_.N && m_.nnzb <= 7L * m_.N;
}

int BiCGStabBackend::solve()
{
    const int rc = solveImpl();
    if (tiled_) tiled_->checkError(m_.stream);
    return rc;
}

int BiCGStabBackend::solveImpl()
{
    if (!prepared_) return 2;   // preconditioner could not be created (NumericalProblem)
    hipStream_t s = m_.stream;
    const long m = (long)m_.bs * m_.rows();
    const double absTol = prm_.lin_abs_tol < 0 ? prm_.newton_tol / 100.0 : prm_.lin_abs_tol;
    iters_ = 0;
    if (prm_.precision == 0) {
        int its = 0;
        int rc = bicgstab<double>(A64_, lu64_.p, dinv64_.p, b64_.p, x64_.p, prm_.lin_tol, absTol,
                                  prm_.lin_max_iter, &its);
        iters_ = its;
        return rc;
    }
    if (prm_.precision == 2) {
        launchCopyConvertD<float>(m, b64_.p, bf_.p, s);
        int its = 0;
        int rc = bicgstab<float>(Af_.p, luf_.p, dinv64_.p, bf_.p, xf_.p, prm_.lin_tol, absTol, prm_.lin_max_iter,
                                 &its);
        iters_ = its;
        OM_HIP(hipMemsetAsync(x64_.p, 0, m * sizeof(double), s));
        launchAddFromFloat(m, xf_.p, x64_.p, s);
        return rc;
    }
    // FP32 inner BiCGStab inside FP64 iterative refinement: the outer residual
    // r = b - A x and its norm are FP64; the inner solve works on FP32 copies
    // of J and of the ILU factors (with FP64 pivots).
    OM_HIP(hipMemsetAsync(x64_.p, 0, m * sizeof(double), s));
    OM_HIP(hipMemcpyAsync(r64_.p, b64_.p, m * sizeof(double), hipMemcpyDeviceToDevice, s));
    launchMaxAbs<double>(m, r64_.p, m_.partial.p, m_.scal.p, s);
    OM_HIP(hipMemcpyAsync(m_.hostScal, m_.scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
    m_.sync();
    const double init = std::max(m_.hostScal[0], std::numeric_limits<double>::min() * 1e10);
    double res = init;
    auto converged = [&](double r) { return r / std::max(1e-20, init) <= prm_.lin_tol || r <= absTol; };
    if (converged(res)) return 0;
    const int maxOuter = 8;
    int total = 0;
    for (int outer = 0; outer < maxOuter; ++outer) {
        launchCopyConvertD<float>(m, r64_.p, bf_.p, s);
        const double innerTol = std::min(0.5, std::max(1e-6, prm_.lin_tol * init / res));
        int its = 0;
        int rc = bicgstab<float>(Af_.p, luf_.p, dinv64_.p, bf_.p, xf_.p, innerTol, absTol,
                                 prm_.lin_max_iter - total, &its);
        total += its;
        if (rc == 2) { iters_ = total; return 2; }
        launchAddFromFloat(m, xf_.p, x64_.p, s);
        launchResidualUpdate((int)m_.rows(), m_.bs, m_.rowPtr.p, m_.colIdx.p, A64_, x64_.p, b64_.p, r64_.p, s);
        launchMaxAbs<double>(m, r64_.p, m_.partial.p, m_.scal.p, s);
        OM_HIP(hipMemcpyAsync(m_.hostScal, m_.scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
        m_.sync();
        res = m_.hostScal[0];
        if (!std::isfinite(res)) { iters_ = total; return 2; }
        if (converged(res)) { iters_ = total; return 0; }
        if (rc == 1 || res > prm_.lin_max_error || total >= prm_.lin_max_iter) { iters_ = total; return 1; }
    }
    iters_ = total;
    return 1;
}

void BiCGStabBackend::downloadIlu(double* hostValues)
{
    if (colored_) throw std::invalid_argument("the two-colour ILU(0) keeps only its pivots");
    lu64_.download(hostValues, (size_t)m_.bs * m_.bs * m_.nnzb, m_.stream);
    m_.sync();
}

const AmgHierarchy& BiCGStabBackend::amg() const
{
    if (!amg_) throw std::invalid_argument("the solver has no AMG (precond != 2)");
    return *amg_;
}

void BiCGStabBackend::downloadAmgLevel(int level, double* hostValues)
{
    const AmgHierarchy& a = amg();
    if (!prepared_) throw std::logic_error("AMG levels before a successful prepare");
    if (level < 0 || level >= a.levels()) throw std::invalid_argument("AMG level out of range");
    a.downloadLevel(level, hostValues, m_.stream);
}

void BiCGStabBackend::applyPrecond(const double* hostD, double* hostV)
{
    const long m = (long)m_.bs * m_.rows();
    DevBuf<double> d, v;
    d.upload(hostD, m, m_.stream);
    v.alloc(m);
    OM_HIP(hipMemsetAsync(v.p, 0, m * sizeof(double), m_.stream));   // Jacobi sweeps start from zero
    if (colored_ && !dia64_.p) {
        // the FP32-inner form keeps only the FP32 structured matrix: the FP64
        // application here reads an FP64 one
        dia64_.alloc(7L * m_.bs * m_.bs * m_.N);
        launchToDia<double>(m_.gx, m_.gy, m_.gz, m_.rowPtr.p, A64_, dia64_.p, m_.stream);
    }
    precApply<double>(lu64_.p, dinv64_.p, d.p, v.p);
    v.download(hostV, m, m_.stream);
    m_.sync();
}

// ---------------------------------------------------------------------------
int BlackOilNewtonMethod::update(const double* delta, int* numSwitched)
{
    hipStream_t s = m_.stream;
    NewtonUpdateParams up{prm_.dp_max_rel, prm_.ds_max, prm_.privar_osc_thresh};
    OM_HIP(hipMemsetAsync(m_.counters.p, 0, 4 * sizeof(int), s));
    launchBoUpdate(m_.fluid, m_.N, m_.pv0.p, m_.mean0.p, m_.wasSw.p, delta, up, m_.counters.p, s);
    // the auxiliary DOFs (newtonmethod.hh:704-708)
    launchAuxUpdate(3 * m_.nAux, m_.aux0.p, delta + 3 * m_.N, m_.counters.p, s);
    int c[4];
    OM_HIP(hipMemcpyAsync(c, m_.counters.p, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
    m_.sync();
    *numSwitched = c[0];
    return c[1] ? 2 : 0;
}

double BlackOilNewtonMethod::suggestTimeStepSize(double oldDt) const
{
    const int tgt = prm_.newton_target_iter;
    if (numIter_ > tgt) {
        const double percent = double(numIter_ - tgt) / tgt;
        return std::max(prm_.min_dt, oldDt / (1.0 + percent));
    }
    const double percent = double(tgt - numIter_) / tgt;
    return std::max(prm_.min_dt, oldDt * (1.0 + percent / 1.2));
}

// NewtonMethod::apply (newtonmethod.hh:228-438) with the black-oil hooks.
bool BlackOilNewtonMethod::apply(double dt, int step, int attempt, std::vector<omodels_iter_log>* log,
                                 omodels_run_stats* st)
{
    using clk = std::chrono::steady_clock;
    numIter_ = 0;
    error_ = 1e100; lastError_ = 1e100;
    auto converged = [&]() { return error_ <= prm_.newton_tol; };
    auto proceed = [&]() {
        if (numIter_ < 1) return true;
        if (converged()) return false;
        if (numIter_ >= prm_.newton_max_iter) return error_ * 4.0 < lastError_;
        return true;
    };
    while (proceed()) {
        auto t0 = clk::now();
        m_.linearizeDomain(dt, numIter_);
        // linearizeAuxiliaryEquations_ (newtonmethod.hh:287, :566-570): a
        // module that fails makes the attempt fail (NumericalProblem)
        if (m_.nAux > 0) {
            if (!m_.auxFn) throw std::invalid_argument("auxiliary DOFs without a module to linearize them");
            if (m_.auxFn(m_.auxUser, m_.auxHandle, numIter_) != 0) {
                numIter_ = 2 * prm_.newton_target_iter;
                return false;
            }
        }
        lastError_ = error_;
        error_ = m_.maxAbsResidual();           // preSolve_ (:572-607)
        auto t1 = clk::now();
        if (st) { st->t_linearize += std::chrono::duration<double>(t1 - t0).count(); st->linearizations++; }
        if (!(error_ <= prm_.newton_max_error)) return false;
        if (!proceed()) {
            if (log) log->push_back({step, attempt, numIter_, 0, 0, 1, error_, dt});
            ++numIter_;
            break;
        }
        solver_.prepare();
        const int rc = solver_.solve();
        auto t2 = clk::now();
        if (st) { st->t_solve += std::chrono::duration<double>(t2 - t1).count(); st->linear_iters += solver_.iterations(); }
        if (rc != 0) {
            if (log) log->push_back({step, attempt, numIter_, solver_.iterations(), 0, 0, error_, dt});
            numIter_ = 2 * prm_.newton_target_iter;     // failed_()
            return false;
        }
        int nsw = 0;
        const int urc = update(solver_.solution(), &nsw);
        auto t3 = clk::now();
        if (st) { st->t_update += std::chrono::duration<double>(t3 - t2).count(); st->newton_iters++; }
        if (log) log->push_back({step, attempt, numIter_, solver_.iterations(), nsw, 1, error_, dt});
        if (prm_.verbose)
            printf("step %d attempt %d iter %d error %.6e lin %d switched %d\n", step, attempt, numIter_, error_,
                   solver_.iterations(), nsw);
        if (urc != 0) { numIter_ = 2 * prm_.newton_target_iter; return false; }
        ++numIter_;
    }
    if (!converged()) { numIter_ = 2 * prm_.newton_target_iter; return false; }
    return true;
}

// Simulator::run / FvBaseProblem::timeIntegration with step halving.
int simulate(Model& m, const omodels_params& prm, int nSteps, double dt0, double* pvHist, double* dtHist,
             std::vector<omodels_iter_log>* log, omodels_run_stats* stats)
{
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    BlackOilNewtonMethod newton(m, prm);
    double dt = dt0;
    m.advanceTimeLevel();
    for (int step = 0; step < nSteps; ++step) {
        bool ok = false;
        for (int attempt = 0; attempt < prm.max_dt_divisions; ++attempt) {
            if (newton.apply(dt, step, attempt, log, stats)) { ok = true; break; }
            if (stats) stats->failed_attempts++;
            m.restoreTimeLevel();
            double ndt = dt / 2.0;
            if (dt < prm.min_dt * (1 + 1e-9)) break;
            if (ndt < prm.min_dt) ndt = prm.min_dt;
            dt = ndt;
        }
        if (!ok) { m.sync(); return 3; }
        if (dtHist) dtHist[step] = dt;
        m.advanceTimeLevel();
        if (pvHist) {
            OM_HIP(hipMemcpyAsync(pvHist + (long)step * 3 * m.N, m.pv0.p, 3 * m.N * sizeof(double),
                                  hipMemcpyDeviceToHost, m.stream));
        }
        dt = std::min(prm.max_dt, newton.suggestTimeStepSize(dt));
    }
    m.sync();
    m.timer.collect();
    if (stats) stats->t_total = std::chrono::duration<double>(clk::now() - t0).count();
    return 0;
}

}  // namespace omodels
