/*
 * blackoil.cpp -- CPU restatement (the oracle) of opm-models' fully-implicit
 * black-oil TPFA path: TpfaLinearizer, BlackOilLocalResidualTPFA,
 * BlackOilIntensiveQuantities, BiCGStab + block ILU0 / block Jacobi,
 * NewtonMethod + BlackOilNewtonMethod and the time-step control.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the parity reference / CPU baseline, never by the
 * product library.  Written as plain, sequential-by-default C++ so that each
 * function reads against the reference file:line it restates.
 *
 * Choices for code that lives outside opm-models (opm-common, dune-istl,
 * opm-simulators; absent here, "parity unpinned" against them):
 *  - DenseAd: value + N derivatives; a/b = a * (1/b), (a/b)' = (a' - (a/b) b') / b
 *    computed with the reciprocal; min/max keep the AD operand at ties.
 *  - LinearMaterial (tests/problems/reservoirproblem.hh:84-100,340-350):
 *    kr_a = max(min(S_a, 1), 0), pc = 0 (pcMinSat = pcMaxSat = 0).  The
 *    relperm hook (multiphasebaseproblem.hh:257) is filled as flow does.
 *  - PVT: piecewise-linear tables in p of 1/B and mu (LiveOil saturated,
 *    DryGas), linear extrapolation outside the table; undersaturated oil:
 *    pb = Rs_sat^{-1}(Rs), 1/Bo = 1/Bo_sat(pb) * (1 + c_o (p - pb)),
 *    mu_o = mu_o_sat(pb).  Water: 1/Bw = 1 + X (1 + X/2), X = c_w (p - p_ref),
 *    constant mu_w (ConstantCompressibilityWater with c_v = 0, Bw_ref = 1).
 *  - Transmissibility: flow's half-transmissibility A K |n.d| / |d|^2 with the
 *    harmonic combination of transfluxmodule.hh:425-456.
 *  - Pressure difference: transfluxmodule.hh:254-303 (no threshold pressure).
 *  - 3x3/2x2 block inverse by the adjugate; det == 0 -> NumericalProblem.
 *
 * Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp).
 */
#include "oracle.h"
#include "ad.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <limits>
#include <set>
#include <stdexcept>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using namespace oracle_ad;
typedef Ev<3> E3;

/* ------------------------------------------------------------------ */
/* Piecewise-linear table (Tabulated1DFunction with extrapolation). */
struct Table {
    std::vector<double> x, y, slope;
    void set(const double* xs, const double* ys, int n)
    {
        x.assign(xs, xs + n); y.assign(ys, ys + n); slope.resize(n - 1);
        for (int i = 0; i + 1 < n; ++i) slope[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
    }
    int seg(double xv) const
    {
        int s = 0;
        for (int i = 1; i + 1 < (int)x.size(); ++i) if (xv >= x[i]) s = i;
        return s;
    }
    double eval(double xv) const { int s = seg(xv); return y[s] + slope[s] * (xv - x[s]); }
    template <int N> Ev<N> eval(const Ev<N>& xv) const
    {
        int s = seg(xv.v); Ev<N> r;
        r.v = y[s] + slope[s] * (xv.v - x[s]);
        for (int i = 0; i < N; ++i) r.d[i] = slope[s] * xv.d[i];
        return r;
    }
};

enum { W = 0, O = 1, G = 2 };                 /* phases (BlackOilFluidSystem) */
enum { CO = 0, CW = 1, CG = 2 };              /* components == equations */
static const int compOfPhase[3] = {CW, CO, CG};
enum { MEAN_SG = 0, MEAN_RS = 1 };            /* GasMeaning used here */

struct Pvt {
    Table rsSat, pOfRs, invBoSat, muoSat, invBg, mug;
    double co, pref, cw, muw, rho[3];
    void init(const ob_problem* p)
    {
        const int n = p->n_oil; const double* t = p->oil_tab;
        rsSat.set(t, t + n, n); pOfRs.set(t + n, t, n);
        invBoSat.set(t, t + 2 * n, n); muoSat.set(t, t + 3 * n, n);
        const int m = p->n_gas; const double* g = p->gas_tab;
        invBg.set(g, g + m, m); mug.set(g, g + 2 * m, m);
        co = p->oil_comp; pref = p->water_pref; cw = p->water_comp; muw = p->water_mu;
        for (int i = 0; i < 3; ++i) rho[i] = p->rho_ref[i];
    }
};

/* Intensive quantities of one cell; T = E3 (own cell, with derivatives) or
 * double (neighbour, value only).  Restates
 * blackoilintensivequantities.hh:179-516 for the Sw/Po/{Sg,Rs} meanings. */
template <class T>
struct IQ {
    T S[3], p[3], kr[3], invB[3], mu[3], lam[3], rho[3], Rs, phi, tm;
};

inline double valueOf(double x) { return x; }
inline double valueOf(const E3& x) { return x.v; }
inline double dmax(double a, double c) { return a >= c ? a : c; }
inline double dmin(double a, double c) { return a <= c ? a : c; }
inline E3 tmin(const E3& a, double c) { return ev_min(a, c); }
inline E3 tmax(const E3& a, double c) { return ev_max(a, c); }
inline double tmin(double a, double c) { return dmin(a, c); }
inline double tmax(double a, double c) { return dmax(a, c); }

template <class T> T makeVar(double x, int idx);
template <> inline E3 makeVar<E3>(double x, int idx) { return E3::var(x, idx); }
template <> inline double makeVar<double>(double x, int) { return x; }
template <class T> inline T tdiv(const T& a, const T& b);
template <> inline E3 tdiv<E3>(const E3& a, const E3& b) { return a / b; }
template <> inline double tdiv<double>(const double& a, const double& b) { return a * (1.0 / b); }

template <class T>
void computeIQ(const Pvt& pvt, const double* pv, int meaning, double poro, IQ<T>& q)
{
    const T Sw = makeVar<T>(pv[0], 0);
    const T Sg = (meaning == MEAN_SG) ? makeVar<T>(pv[2], 2) : T(0.0);
    const T So = (1.0 - Sw) - Sg;
    q.S[W] = Sw; q.S[O] = So; q.S[G] = Sg;
    const T po = makeVar<T>(pv[1], 1);
    /* LinearMaterial: pc == 0, hence p_a = po + (pc_a - pc_o) = po */
    for (int a = 0; a < 3; ++a) q.p[a] = po;
    for (int a = 0; a < 3; ++a) q.kr[a] = tmax(tmin(q.S[a], 1.0), 0.0);
    /* dissolved gas: blackoilintensivequantities.hh:309-323 */
    if (meaning == MEAN_RS) q.Rs = makeVar<T>(pv[2], 2);
    else q.Rs = pvt.rsSat.eval(po);
    /* water: ConstantCompressibilityWater */
    {
        T X = (q.p[W] - pvt.pref) * pvt.cw;
        T u = X * 0.5 + 1.0;
        q.invB[W] = X * u + 1.0;
        q.mu[W] = T(pvt.muw);
    }
    /* oil: LiveOil */
    if (meaning == MEAN_SG) {
        q.invB[O] = pvt.invBoSat.eval(q.p[O]);
        q.mu[O] = pvt.muoSat.eval(q.p[O]);
    } else {
        T pb = pvt.pOfRs.eval(q.Rs);
        T f = (q.p[O] - pb) * pvt.co + 1.0;
        q.invB[O] = pvt.invBoSat.eval(pb) * f;
        q.mu[O] = pvt.muoSat.eval(pb);
    }
    /* gas: DryGas */
    q.invB[G] = pvt.invBg.eval(q.p[G]);
    q.mu[G] = pvt.mug.eval(q.p[G]);
    /* mobilities: blackoilintensivequantities.hh:366-397 */
    for (int a = 0; a < 3; ++a) q.lam[a] = tdiv<T>(q.kr[a], q.mu[a]);
    /* densities: blackoilintensivequantities.hh:400-442 */
    q.rho[W] = q.invB[W] * pvt.rho[W];
    q.rho[G] = q.invB[G] * pvt.rho[G];
    q.rho[O] = q.invB[O] * pvt.rho[O] + (q.invB[O] * q.Rs) * pvt.rho[G];
    /* porosity; without rock terms the problem's defaults (compressibility
     * 0, multipliers 1, blackoilproblem.hh:131-181) */
    q.phi = T(poro);
    q.tm = T(1.0);
}

/* Rock compressibility and compaction (blackoilintensivequantities.hh:447-
 * 480): porosity *= 1 + x + 0.5 x^2 with x = c (p_o - p_ref) where c > 0,
 * then *= rockCompPoroMultiplier(p_o); rockCompTransMultiplier(p_o) for the
 * face term. */
struct Rock {
    const double* comp = nullptr;
    const double* pref = nullptr;
    bool tab = false;
    Table poroMult, transMult;
};

inline bool rockOn(const ob_problem* p) { return p->rock_comp || p->n_rock > 0; }
inline Rock makeRock(const ob_problem* p)
{
    Rock r;
    const long N = (long)p->nx * p->ny * p->nz;
    if (p->rock_comp)
        for (long I = 0; I < N; ++I)
            if (!(p->rock_comp[I] >= 0.0) || !std::isfinite(p->rock_comp[I]))
                throw std::invalid_argument("rock compressibilities must be finite and >= 0");
    if (p->rock_pref)
        for (long I = 0; I < N; ++I)
            if (!std::isfinite(p->rock_pref[I])) throw std::invalid_argument("rock reference pressures must be finite");
    r.comp = p->rock_comp;
    r.pref = p->rock_pref;
    if (p->n_rock > 0) {
        const int n = p->n_rock;
        if (n < 2 || !p->rock_tab) throw std::invalid_argument("rock table needs at least two nodes");
        for (int k = 0; k < 2 * n; ++k)
            if (!(p->rock_tab[n + k] > 0.0) || !std::isfinite(p->rock_tab[n + k]))
                throw std::invalid_argument("rock multipliers must be finite and > 0");
        for (int k = 1; k < n; ++k)
            if (!(p->rock_tab[k] > p->rock_tab[k - 1])) throw std::invalid_argument("rock table pressures must increase");
        r.tab = true;
        r.poroMult.set(p->rock_tab, p->rock_tab + n, n);
        r.transMult.set(p->rock_tab, p->rock_tab + 2 * n, n);
    }
    return r;
}

template <class T>
void applyRock(const Rock& r, long cell, IQ<T>& q)
{
    const double c = r.comp ? r.comp[cell] : 0.0;
    if (c > 0.0) {
        const double pref = r.pref ? r.pref[cell] : 1e5;
        const T x = (q.p[O] - pref) * c;
        q.phi = q.phi * ((x + 1.0) + (x * 0.5) * x);
    }
    if (r.tab) {
        q.phi = q.phi * r.poroMult.eval(q.p[O]);
        q.tm = r.transMult.eval(q.p[O]);
    }
}

/* ------------------------------------------------------------------ */
/* Grid helpers (YaspGrid order, EcfvStencil intersection order). */
struct Grid {
    int nx, ny, nz; long N;
    double dx, dy, dz, g;
    const ob_problem* P;
    explicit Grid(const ob_problem* p)
        : nx(p->nx), ny(p->ny), nz(p->nz), N((long)p->nx * p->ny * p->nz),
          dx(p->dx), dy(p->dy), dz(p->dz), g(p->gravity), P(p) {}
    /* dofTotalVolume and dofCenterDepth: the caller's, or the structured grid's */
    double volume(long I) const { return P->volume ? P->volume[I] : dx * dy * dz; }
    double zc(long I) const
    {
        if (P->depth) return P->depth[I];
        long k = I / ((long)nx * ny);
        return dz * (k + 0.5);
    }
    /* thresholdPressure(I, J), 0 without the caller's */
    double thpres(long I, long J, int dir) const
    {
        return P->thpres ? P->thpres[(long)(dir / 2) * N + (I < J ? I : J)] : 0.0;
    }
    /* neighbour of I across intersection `dir` (0:-x 1:+x 2:-y 3:+y 4:-z 5:+z), -1 if boundary */
    long nb(long I, int dir) const
    {
        long i = I % nx, j = (I / nx) % ny, k = I / ((long)nx * ny);
        switch (dir) {
        case 0: return i > 0 ? I - 1 : -1;
        case 1: return i + 1 < nx ? I + 1 : -1;
        case 2: return j > 0 ? I - nx : -1;
        case 3: return j + 1 < ny ? I + nx : -1;
        case 4: return k > 0 ? I - (long)nx * ny : -1;
        default: return k + 1 < nz ? I + (long)nx * ny : -1;
        }
    }
    double area(int dir) const
    {
        int ax = dir / 2;
        return ax == 0 ? dy * dz : (ax == 1 ? dx * dz : dx * dy);
    }
    double h(int dir) const { int ax = dir / 2; return ax == 0 ? dx : (ax == 1 ? dy : dz); }
    /* transfluxmodule.hh:425-456 with flow's area factor */
    double trans(long I, long J, int dir) const
    {
        const int ax = dir / 2;
        if (P->face_trans) return P->face_trans[(long)ax * N + (I < J ? I : J)];
        const double A = area(dir), ndot = 0.5 * h(dir), dist2 = ndot * ndot;
        const double K0 = P->perm[(long)ax * N + I], K1 = P->perm[(long)ax * N + J];
        const double T0 = A * K0 * ndot / dist2;
        const double T1 = A * K1 * ndot / dist2;
        return T0 * T1 / (T0 + T1);
    }
};

struct NbInfo { long J; int off; double trans, area, dZg; int dir; double diffusivity, dispersivity, thpres; };

struct Pattern {
    std::vector<int32_t> rowPtr, colIdx, diagOff, nbPtr;
    std::vector<NbInfo> nb;
};

/* TpfaLinearizer::createMatrix_ (tpfalinearizer.hh:414-525) with
 * IstlSparseMatrixAdapter::reserve (istlsparsematrixadapter.hh:76-101). */
void buildPattern(const Grid& G, Pattern& P)
{
    const long N = G.N;
    /* the auxiliary module's degrees of freedom follow the cells
     * (numTotalDof = numGridDof + the modules' numDofs, setDofOffset) */
    const ob_problem* pp = G.P;
    const long nAux = pp->n_aux;
    if (nAux < 0 || pp->n_aux_links < 0) throw std::invalid_argument("auxiliary: counts must be >= 0");
    if (nAux > 0 && (pp->n_aux_links == 0 || !pp->aux_row || !pp->aux_col))
        throw std::invalid_argument("auxiliary: the module's pattern entries");
    const long R = N + nAux;
    std::vector<std::set<long>> sp(R);
    P.nbPtr.assign(N + 1, 0);
    P.nb.clear();
    for (long I = 0; I < N; ++I) {
        sp[I].insert(I);
        for (int dir = 0; dir < 6; ++dir) {
            long J = G.nb(I, dir);
            if (J < 0) continue;
            sp[I].insert(J);
            NbInfo n;
            n.J = J; n.off = -1; n.trans = G.trans(I, J, dir); n.area = G.area(dir);
            n.dZg = (G.zc(I) - G.zc(J)) * G.g; n.dir = dir; n.diffusivity = 0.0; n.dispersivity = 0.0;
            n.thpres = G.thpres(I, J, dir);
            P.nb.push_back(n);
        }
        P.nbPtr[I + 1] = (int32_t)P.nb.size();
    }
    /* addNeighbors of the auxiliary module (tpfalinearizer.hh:503-507) */
    for (long k = 0; nAux > 0 && k < pp->n_aux_links; ++k) {
        const long r = pp->aux_row[k], c = pp->aux_col[k];
        if (r < 0 || r >= R || c < 0 || c >= R) throw std::invalid_argument("auxiliary: entry out of range");
        sp[r].insert(c);
    }
    for (long r = N; r < R; ++r)
        if (!sp[r].count(r)) throw std::invalid_argument("auxiliary: a row without its diagonal block");
    P.rowPtr.assign(R + 1, 0);
    for (long I = 0; I < R; ++I) P.rowPtr[I + 1] = P.rowPtr[I] + (int32_t)sp[I].size();
    P.colIdx.resize(P.rowPtr[R]);
    for (long I = 0; I < R; ++I) {
        int k = P.rowPtr[I];
        for (long c : sp[I]) P.colIdx[k++] = (int32_t)c;
    }
    auto blockAddress = [&](long row, long col) -> int32_t {
        const int32_t* b = P.colIdx.data() + P.rowPtr[row];
        const int32_t* e = P.colIdx.data() + P.rowPtr[row + 1];
        const int32_t* it = std::lower_bound(b, e, (int32_t)col);
        if (it == e || *it != col) throw std::logic_error("missing block");
        return (int32_t)(it - P.colIdx.data());
    };
    P.diagOff.resize(R);
    for (long I = 0; I < N; ++I) {
        P.diagOff[I] = blockAddress(I, I);
        for (int k = P.nbPtr[I]; k < P.nbPtr[I + 1]; ++k) P.nb[k].off = blockAddress(P.nb[k].J, I);
    }
    for (long r = N; r < R; ++r) P.diagOff[r] = blockAddress(r, r);
}

/* BlackOilLocalResidualTPFA::computeStorage (blackoillocalresidualtpfa.hh:151-227)
 * incl. adaptMassConservationQuantities_ (:824-851). */
template <class T>
void storageOf(const Pvt& pvt, const IQ<T>& q, T st[3])
{
    for (int e = 0; e < 3; ++e) st[e] = T(0.0);
    for (int a = 0; a < 3; ++a) {
        T sv = (q.S[a] * q.invB[a]) * q.phi;
        st[compOfPhase[a]] = st[compOfPhase[a]] + sv;
        if (a == O) st[CG] = st[CG] + q.Rs * sv;
    }
    st[CW] = st[CW] * pvt.rho[W];
    st[CG] = st[CG] * pvt.rho[G];
    st[CO] = st[CO] * pvt.rho[O];
}

/* Constants of the diffusive face term.  toFrac is toFractionGasOil
 * (blackoildiffusionmodule.hh:242-248): rho_o M_g / (rho_g M_o), with the
 * molar masses replaced by 1 unless the mole-fraction form is asked for.
 * effOil/effGas: the effective coefficients of the oil and the gas component
 * in the oil phase, the arithmetic mean of two cells' constant coefficients
 * (:568-590; tortuosity is disabled for the black-oil model, :384-394). */
struct Diffusion {
    double toFrac, effOil, effGas;
};

/* Dispersion in the oil phase: toMassFrac = rho_o / rho_g
 * (toMassFractionGasOil, blackoildispersionmodule.hh:258-262); nvI / nvJ the
 * two cells' oil normVelocityCell of the previous linearization. */
struct Dispersion {
    double toMassFrac;
};

/* calculateFluxes_ (blackoillocalresidualtpfa.hh:325-504) with the
 * pressure difference of transfluxmodule.hh:254-303 and evalPhaseFluxes_
 * (:752-811), followed by the caller's multiplication with the face area
 * (tpfalinearizer.hh:758). */
/* darcyOut (optional): each phase's darcyFlux.value() * faceArea at its
 * component index, the FLORES values (blackoillocalresidualtpfa.hh:376-385) */
void faceFlux(const Pvt& pvt, const IQ<E3>& in, const IQ<double>& ex, long I, long J,
              const NbInfo& nbi, double Vin, double Vex, E3 flux[3], double* darcyOut = nullptr,
              const Diffusion* dif = nullptr, const Dispersion* dsp = nullptr, double nvI = 0.0,
              double nvJ = 0.0, bool tmOn = false, const double* dirMob = nullptr, long N = 0)
{
    /* the mean of the two cells' rockCompTransMultiplier
     * (blackoillocalresidualtpfa.hh:372-381) */
    const E3 transMult = (in.tm + ex.tm) / 2.0;
    for (int e = 0; e < 3; ++e) flux[e] = E3(0.0);
    if (darcyOut) for (int e = 0; e < 3; ++e) darcyOut[e] = 0.0;
    for (int a = 0; a < 3; ++a) {
        if (in.lam[a].v <= 0.0 && ex.lam[a] <= 0.0) continue;
        E3 rhoAvg = (in.rho[a] + ex.rho[a]) * 0.5;
        E3 pEx = rhoAvg * nbi.dZg + ex.p[a];
        E3 dp = pEx - in.p[a];
        /* threshold pressure: the form of the reference's solvent flux
         * (blackoilsolventmodules.hh:1484-1492), applied only where a face
         * has one; below it the face carries no flow */
        if (nbi.thpres > 0.0) {
            if (std::fabs(dp.v) > nbi.thpres) dp = dp.v < 0.0 ? dp + nbi.thpres : dp - nbi.thpres;
            else dp = E3(0.0);
        }
        bool upIn;
        if (dp.v > 0.0) upIn = false;
        else if (dp.v < 0.0) upIn = true;
        else if (Vin > Vex) upIn = true;
        else if (Vin < Vex) upIn = false;
        else upIn = (I < J);
        const double tf = (-nbi.trans) / nbi.area;
        E3 darcy, svf;
        /* the upwind cell's mobility along the face's axis
         * (BlackOilIntensiveQuantities::mobility(phaseIdx, facedir),
         * blackoilintensivequantities.hh:530-550): kr_dir = m_dir kr */
        const long dmAt = (3L * (nbi.dir / 2) + a) * N;
        if (upIn) {
            const E3 lam = dirMob ? in.lam[a] * dirMob[dmAt + I] : in.lam[a];
            darcy = tmOn ? ((dp * lam) * transMult) * tf : (dp * lam) * tf;
            svf = in.invB[a] * darcy;
        } else {
            const double lam = dirMob ? ex.lam[a] * dirMob[dmAt + J] : ex.lam[a];
            darcy = tmOn ? dp * ((transMult * lam) * tf) : dp * (lam * tf);
            svf = darcy * ex.invB[a];
        }
        flux[compOfPhase[a]] = flux[compOfPhase[a]] + svf * pvt.rho[a];
        if (a == O) {
            E3 gs = upIn ? (in.Rs * svf) : (svf * ex.Rs);
            flux[CG] = flux[CG] + gs * pvt.rho[G];
        }
        if (darcyOut) darcyOut[compOfPhase[a]] = darcy.v * nbi.area;
    }
    /* BlackOilDiffusionModule::addDiffusiveFlux (blackoildiffusionmodule.hh:168-238),
     * added after the phase loop (blackoillocalresidualtpfa.hh:469-479) with the
     * face's diffusivity per area.  Only the oil phase carries a term here:
     * water is skipped (no gas in water, :180-183) and the gas phase, without
     * vaporized oil, has diffR = 0, so its terms are exact zeros (:205-209).
     * The terms enter in surface-volume units, as the reference adds them. */
    if (dif) {
        E3 bSAvg = in.S[O] * in.invB[O];
        bSAvg = bSAvg + ex.S[O] * ex.invB[O];
        bSAvg = bSAvg / 2.0;
        if (!(bSAvg.v < 1.0e-6)) {
            const E3 rsAvg = (in.Rs + ex.Rs) / 2.0;
            const E3 convFactor = E3(1.0) / (rsAvg + dif->toFrac);
            const E3 diffR = in.Rs - ex.Rs;
            const double tmpDiff = nbi.diffusivity / nbi.area;
            /* solvent (oil in oil) and solute (gas in oil) */
            flux[CO] = flux[CO] + (((-bSAvg * convFactor) * diffR) * tmpDiff) * dif->effOil;
            flux[CG] = flux[CG] + (((bSAvg * diffR) * convFactor) * tmpDiff) * dif->effGas;
        }
    }
    /* BlackOilDispersionModule::addDispersiveFlux (blackoildispersionmodule.hh:176-254)
     * after the diffusive term (blackoillocalresidualtpfa.hh:481-493): the oil
     * phase only (no gas in water; the gas phase is disabled, :193-197), with
     * the mean 1/Bo, the mean of the two cells' normVelocityCell (:484-506)
     * and the face's dispersivity per area. */
    if (dsp) {
        E3 bAvg = in.invB[O];
        bAvg = bAvg + ex.invB[O];
        bAvg = bAvg / 2.0;
        const E3 rsAvg = (in.Rs + ex.Rs) / 2.0;
        const E3 convFactor = E3(1.0) / (rsAvg + dsp->toMassFrac);
        const E3 diffR = in.Rs - ex.Rs;
        const double nvAvg = 0.5 * (nvI + nvJ);
        const double tmpDisp = nbi.dispersivity / nbi.area;
        flux[CO] = flux[CO] + ((((-bAvg * nvAvg) * convFactor) * tmpDisp) * diffR);
        flux[CG] = flux[CG] + ((((bAvg * nvAvg) * convFactor) * tmpDisp) * diffR);
    }
    for (int e = 0; e < 3; ++e) flux[e] = flux[e] * nbi.area;
}

/* ReservoirProblem-style source(): injectors add water mass, producers are
 * bottom-hole-pressure controlled.  Rate per unit volume. */
void sourceOf(const Pvt& pvt, const IQ<E3>& q, int type, double a, double b, double V, E3 src[3])
{
    for (int e = 0; e < 3; ++e) src[e] = E3(0.0);
    if (type == 0) {
        src[CW] = E3(a / V);
        return;
    }
    E3 drive = q.p[O] - b;
    if (drive.v <= 0.0) return;
    for (int ph = 0; ph < 3; ++ph) {
        E3 qs = ((q.lam[ph] * q.invB[ph]) * drive) * (-a);
        src[compOfPhase[ph]] = src[compOfPhase[ph]] + (qs * pvt.rho[ph]) * (1.0 / V);
        if (ph == O) src[CG] = src[CG] + ((q.Rs * qs) * pvt.rho[G]) * (1.0 / V);
    }
}

struct Problem {
    const ob_problem* P;
    Grid G;
    Pvt pvt;
    Pattern pat;
    /* the problem's source terms per cell: the first source of the cell and
     * the next source of the same cell (-1: none), in the list's order */
    std::vector<int> srcOfCell, srcNext;
    bool hasRock = false;
    Rock rock;
    /* the intensive quantities of cell I with the rock terms */
    template <class T>
    void iq(const double* pv, const uint8_t* mean, long I, IQ<T>& q) const
    {
        computeIQ<T>(pvt, pv + 3 * I, mean[I], P->poro[I], q);
        if (hasRock) applyRock(rock, I, q);
    }
    bool hasDiff = false;
    Diffusion dif{};
    bool hasDisp = false;
    Dispersion dsp{};
    explicit Problem(const ob_problem* p, const ob_diffusion* d = nullptr, const ob_dispersion* e = nullptr)
        : P(p), G(p)
    {
        if (p->face_trans) checkNonNegative(p->face_trans, 3 * G.N, "transmissibilities must be finite and >= 0");
        if (p->dir_mob) checkNonNegative(p->dir_mob, 9 * G.N, "directional mobility multipliers must be finite and >= 0");
        if (p->thpres) checkNonNegative(p->thpres, 3 * G.N, "threshold pressures must be finite and >= 0");
        if (p->depth)
            for (long I = 0; I < G.N; ++I)
                if (!std::isfinite(p->depth[I])) throw std::invalid_argument("depths must be finite");
        if (p->volume)
            for (long I = 0; I < G.N; ++I)
                if (!(p->volume[I] > 0.0) || !std::isfinite(p->volume[I]))
                    throw std::invalid_argument("cell volumes must be finite and > 0");
        pvt.init(p);
        hasRock = rockOn(p);
        if (hasRock) rock = makeRock(p);
        buildPattern(G, pat);
        srcOfCell.assign(G.N, -1);
        srcNext.assign(p->n_src, -1);
        std::vector<int> last(G.N, -1);
        for (int s = 0; s < p->n_src; ++s) {
            const long c = p->src_cell[s];
            if (c < 0 || c >= G.N) throw std::invalid_argument("source cell out of range");
            if (last[c] < 0) srcOfCell[c] = s;
            else srcNext[last[c]] = s;
            last[c] = s;
        }
        if (d) setDiffusion(*d);
        if (e) setDispersion(*e);
        buildBoundary();
    }
    /* the boundary faces with a condition, in cell and then face order (the
     * order TpfaLinearizer::createMatrix_ collects them, tpfalinearizer.hh:
     * 476-497) */
    struct Bc {
        long cell;
        int face, type;
        double rate[3], state[3], trans, zf;
        uint8_t meaning;
    };
    std::vector<Bc> bcs;
    void buildBoundary()
    {
        const ob_problem* p = P;
        bcs.clear();
        if (p->n_bc <= 0) return;
        if (!p->bc_cell || !p->bc_face || !p->bc_type) throw std::invalid_argument("boundary: cells, faces and types");
        for (int b = 0; b < p->n_bc; ++b) {
            Bc c{};
            c.cell = p->bc_cell[b];
            c.face = p->bc_face[b];
            c.type = p->bc_type[b];
            if (c.cell < 0 || c.cell >= G.N || c.face < 0 || c.face > 5)
                throw std::invalid_argument("boundary: cell or face out of range");
            if (G.nb(c.cell, c.face) >= 0) throw std::invalid_argument("boundary: the face has a neighbour");
            if (c.type < 0 || c.type > 3) throw std::invalid_argument("boundary: type must be 0..3");
            if (c.type == 1) {
                if (!p->bc_rate) throw std::invalid_argument("boundary: rate faces need rates");
                for (int e = 0; e < 3; ++e) {
                    c.rate[e] = p->bc_rate[3 * b + e];
                    if (!std::isfinite(c.rate[e])) throw std::invalid_argument("boundary: rates must be finite");
                }
            }
            if (c.type >= 2) {
                if (!p->bc_state || !p->bc_meaning) throw std::invalid_argument("boundary: free faces need a state");
                for (int e = 0; e < 3; ++e) {
                    c.state[e] = p->bc_state[3 * b + e];
                    if (!std::isfinite(c.state[e])) throw std::invalid_argument("boundary: states must be finite");
                }
                c.meaning = p->bc_meaning[b];
                if (c.meaning != MEAN_SG && c.meaning != MEAN_RS) throw std::invalid_argument("boundary: meaning");
            }
            const int ax = c.face / 2;
            if (p->bc_trans) {
                c.trans = p->bc_trans[b];
                if (!(c.trans >= 0.0) || !std::isfinite(c.trans))
                    throw std::invalid_argument("boundary: transmissibilities must be finite and >= 0");
            } else {
                /* the cell's half of a face's transmissibility (transfluxmodule.hh:457-479) */
                const double A = G.area(c.face), ndot = 0.5 * G.h(c.face), dist2 = ndot * ndot;
                c.trans = A * p->perm[(long)ax * G.N + c.cell] * ndot / dist2;
            }
            if (p->bc_zface) {
                c.zf = p->bc_zface[b];
                if (!std::isfinite(c.zf)) throw std::invalid_argument("boundary: face depths must be finite");
            } else {
                c.zf = ax < 2 ? G.zc(c.cell) : (c.face == 4 ? G.zc(c.cell) - 0.5 * G.dz : G.zc(c.cell) + 0.5 * G.dz);
            }
            bcs.push_back(c);
        }
        std::stable_sort(bcs.begin(), bcs.end(), [](const Bc& a, const Bc& b) {
            return a.cell != b.cell ? a.cell < b.cell : a.face < b.face;
        });
        for (size_t k = 1; k < bcs.size(); ++k)
            if (bcs[k].cell == bcs[k - 1].cell && bcs[k].face == bcs[k - 1].face)
                throw std::invalid_argument("boundary: a face given twice");
    }
    /* the problem's dispersivity(I, J) per face (tpfalinearizer.hh:465-467):
     * given per face, or the harmonic mean of A E |n.d| / |d|^2 of the two
     * cells' dispersion coefficients E */
    /* harmonic combination of two half-face values; a face between two cells
     * without porosity (diffusion) or without a dispersion coefficient has
     * none: 0 rather than 0/0 */
    static double harmonicFace(double a, double b) { return a + b == 0.0 ? 0.0 : a * b / (a + b); }
    static void checkNonNegative(const double* v, long n, const char* what)
    {
        for (long i = 0; i < n; ++i)
            if (!(v[i] >= 0.0) || !std::isfinite(v[i])) throw std::invalid_argument(what);
    }
    void setDispersion(const ob_dispersion& e)
    {
        if (!e.face_dispersivity && !e.cell_dispersion)
            throw std::invalid_argument("dispersion needs face dispersivities or cell coefficients");
        if (e.face_dispersivity) checkNonNegative(e.face_dispersivity, 3 * G.N, "face dispersivities must be finite and >= 0");
        else checkNonNegative(e.cell_dispersion, G.N, "dispersion coefficients must be finite and >= 0");
        dsp.toMassFrac = pvt.rho[O] / pvt.rho[::G];
        hasDisp = true;
        for (long I = 0; I < G.N; ++I)
            for (int k = pat.nbPtr[I]; k < pat.nbPtr[I + 1]; ++k) {
                NbInfo& n = pat.nb[k];
                const int ax = n.dir / 2;
                const long lo = I < n.J ? I : n.J, hi = I < n.J ? n.J : I;
                if (e.face_dispersivity) {
                    n.dispersivity = e.face_dispersivity[(long)ax * G.N + lo];
                } else {
                    const double A = G.area(n.dir), ndot = 0.5 * G.h(n.dir), dist2 = ndot * ndot;
                    const double E0 = A * e.cell_dispersion[lo] * ndot / dist2;
                    const double E1 = A * e.cell_dispersion[hi] * ndot / dist2;
                    n.dispersivity = harmonicFace(E0, E1);
                }
            }
    }
    const Dispersion* dispersion() const { return hasDisp ? &dsp : nullptr; }
    /* the problem's diffusivity(I, J) per face (tpfalinearizer.hh:462-464):
     * given per face, or the harmonic mean of the half-face values
     * A phi |n.d| / |d|^2 (the transmissibility form with porosity in place
     * of permeability) */
    void setDiffusion(const ob_diffusion& d)
    {
        if (!(d.oil_in_oil >= 0.0) || !(d.gas_in_oil >= 0.0))
            throw std::invalid_argument("diffusion coefficients must be >= 0");
        const bool mole = d.mw_oil > 0.0 && d.mw_gas > 0.0;
        const double mO = mole ? d.mw_oil : 1.0, mG = mole ? d.mw_gas : 1.0;
        dif.toFrac = pvt.rho[O] * mG / (pvt.rho[::G] * mO);
        dif.effOil = 0.5 * (d.oil_in_oil + d.oil_in_oil);
        dif.effGas = 0.5 * (d.gas_in_oil + d.gas_in_oil);
        if (d.face_diffusivity) checkNonNegative(d.face_diffusivity, 3 * G.N, "face diffusivities must be finite and >= 0");
        hasDiff = true;
        for (long I = 0; I < G.N; ++I)
            for (int k = pat.nbPtr[I]; k < pat.nbPtr[I + 1]; ++k) {
                NbInfo& n = pat.nb[k];
                const int ax = n.dir / 2;
                const long lo = I < n.J ? I : n.J, hi = I < n.J ? n.J : I;
                if (d.face_diffusivity) {
                    n.diffusivity = d.face_diffusivity[(long)ax * G.N + lo];
                } else {
                    const double A = G.area(n.dir), ndot = 0.5 * G.h(n.dir), dist2 = ndot * ndot;
                    const double D0 = A * P->poro[lo] * ndot / dist2;
                    const double D1 = A * P->poro[hi] * ndot / dist2;
                    n.diffusivity = harmonicFace(D0, D1);
                }
            }
    }
    const Diffusion* diffusion() const { return hasDiff ? &dif : nullptr; }
};

void computeStorageAll(const Problem& pr, const double* pv, const uint8_t* mean, double* st)
{
    for (long I = 0; I < pr.G.N; ++I) {
        IQ<double> q;
        pr.iq<double>(pv, mean, I, q);
        double s[3];
        storageOf<double>(pr.pvt, q, s);
        for (int e = 0; e < 3; ++e) st[3 * I + e] = s[e];
    }
}

/* TpfaLinearizer::linearize_ (tpfalinearizer.hh:711-866): resetSystem_ then
 * one pass over cells, each cell owning r[I], A[I][I] and the column blocks
 * A[J][I] of its neighbours. */
/* BlackOilLocalResidualTPFA::computeBoundaryFlux (blackoillocalresidualtpfa.hh:
 * 507-620) times the face area (tpfalinearizer.hh:849-865).  Rate faces:
 * the mass rates per area (setMassRate without the surface-volume
 * conversion, blackoilratevector.hh:92-117).  Free faces: per phase the
 * gradient of the trans flux module's boundary faces (transfluxmodule.hh:
 * 323-409) per area: the boundary pressure plus the mean density times the
 * depth difference to the face, the upstream side by that difference (ties:
 * the cell), the upstream mobility (the boundary's from its own saturations
 * and viscosities); the flux leaves with the cell's 1/B and Rs when the
 * boundary pressure is below the cell's, enters with the boundary's when it
 * is above, nothing when equal; evalPhaseFluxes_ (:753-811) with the
 * reference densities, then adaptMassConservationQuantities_ (:825-851)
 * multiplies by them once more, as the reference does. */
void boundaryFlux(const Pvt& pvt, const IQ<E3>& qi, const Problem::Bc& bc, double area, double zIn, double g,
                  E3 out[3])
{
    for (int e = 0; e < 3; ++e) out[e] = E3(0.0);
    if (bc.type == 1) {
        for (int e = 0; e < 3; ++e) out[e] = E3(bc.rate[e] * area);
        return;
    }
    IQ<double> qb;
    computeIQ<double>(pvt, bc.state, bc.meaning, 0.0, qb);
    const double dZg = (zIn - bc.zf) * g;
    const double tf = (-bc.trans) / area;
    E3 bdy[3] = {E3(0.0), E3(0.0), E3(0.0)};
    for (int a = 0; a < 3; ++a) {
        E3 tmp[3] = {E3(0.0), E3(0.0), E3(0.0)};
        const E3 rhoAvg = (qi.rho[a] + qb.rho[a]) * 0.5;
        const E3 pEx = rhoAvg * dZg + qb.p[a];
        const E3 dp = pEx - qi.p[a];
        const E3 darcy = dp.v > 0.0 ? (dp * qb.lam[a]) * tf : (dp * qi.lam[a]) * tf;
        const double pB = qb.p[a], pIn = qi.p[a].v;
        const int c = compOfPhase[a];
        if (pB < pIn) {
            const E3 svf = qi.invB[a] * darcy;
            tmp[c] = tmp[c] + svf * pvt.rho[a];
            if (a == O) tmp[CG] = tmp[CG] + (qi.Rs * svf) * pvt.rho[G];
        } else if (pB > pIn) {
            const E3 svf = darcy * qb.invB[a];
            tmp[c] = tmp[c] + svf * pvt.rho[a];
            if (a == O) tmp[CG] = tmp[CG] + (svf * qb.Rs) * pvt.rho[G];
        }
        for (int e = 0; e < 3; ++e) bdy[e] = bdy[e] + tmp[e];
    }
    bdy[CW] = bdy[CW] * pvt.rho[W];
    bdy[CG] = bdy[CG] * pvt.rho[G];
    bdy[CO] = bdy[CO] * pvt.rho[O];
    for (int e = 0; e < 3; ++e) out[e] = bdy[e] * area;
}

/* With dispersion, nvIn holds each cell's oil normVelocityCell from the
 * previous linearization (zeros before the first, createFlows_,
 * tpfalinearizer.hh:571-619) and nvOut receives this linearization's: the
 * largest |darcy[oil] / faceArea| over the cell's faces (:759-761 with
 * blackoildispersionmodule.hh:355-383). */
void linearize(const Problem& pr, const double* pv, const uint8_t* mean, const double* st1,
               double dt, double* res, double* jac, int nthreads, const double* nvIn = nullptr,
               double* nvOut = nullptr, const uint8_t* domain = nullptr)
{
    const long N = pr.G.N;
    /* resetSystem_ zeroes every row, the auxiliary ones included; on a
     * sub-domain resetSystem_(domain) only the domain's rows (:370-379) */
    const long R = (long)pr.pat.rowPtr.size() - 1;
    const long nnzb = pr.pat.rowPtr[R];
    if (!domain) {
        std::fill(res, res + 3 * R, 0.0);
        std::fill(jac, jac + 9 * nnzb, 0.0);
    } else {
        for (long I = 0; I < N; ++I) {
            if (!domain[I]) continue;
            std::fill(res + 3 * I, res + 3 * I + 3, 0.0);
            std::fill(jac + 9L * pr.pat.rowPtr[I], jac + 9L * pr.pat.rowPtr[I + 1], 0.0);
        }
    }
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (long I = 0; I < N; ++I) {
        if (domain && !domain[I]) continue;     /* linearize_(domain): the domain's cells (:733-735) */
        IQ<E3> qi;
        pr.iq<E3>(pv, mean, I, qi);
        double* r = res + 3 * I;
        double* D = jac + 9L * pr.pat.diagOff[I];
        const double V = pr.G.volume(I);
        if (pr.hasDisp) nvOut[I] = 0.0;
        for (int k = pr.pat.nbPtr[I]; k < pr.pat.nbPtr[I + 1]; ++k) {
            const NbInfo& nbi = pr.pat.nb[k];
            IQ<double> qj;
            pr.iq<double>(pv, mean, nbi.J, qj);
            E3 f[3];
            if (pr.hasDisp) {
                double darcy[3];
                faceFlux(pr.pvt, qi, qj, I, nbi.J, nbi, V, pr.G.volume(nbi.J), f, darcy, pr.diffusion(), &pr.dsp,
                         nvIn[I], nvIn[nbi.J], pr.rock.tab, pr.P->dir_mob, N);
                const double v = std::fabs(darcy[CO] / nbi.area);
                nvOut[I] = nvOut[I] < v ? v : nvOut[I];
            } else {
                faceFlux(pr.pvt, qi, qj, I, nbi.J, nbi, V, pr.G.volume(nbi.J), f, nullptr, pr.diffusion(), nullptr,
                         0.0, 0.0, pr.rock.tab, pr.P->dir_mob, N);
            }
            double* O_ = jac + 9L * nbi.off;
            for (int e = 0; e < 3; ++e) {
                r[e] += f[e].v;
                for (int pvi = 0; pvi < 3; ++pvi) {
                    D[3 * e + pvi] += f[e].d[pvi];
                    O_[3 * e + pvi] += -f[e].d[pvi];
                }
            }
        }
        /* accumulation (:776-825), storage cache of the step start */
        {
            E3 s[3];
            storageOf<E3>(pr.pvt, qi, s);
            const double storefac = V / dt;
            for (int e = 0; e < 3; ++e) {
                double rv = s[e].v;
                rv -= st1[3 * I + e];
                rv *= storefac;
                r[e] += rv;
                for (int pvi = 0; pvi < 3; ++pvi) D[3 * e + pvi] += s[e].d[pvi] * storefac;
            }
        }
        /* cell-wise sources (:827-841) */
        /* problem.source(globI) is the cell's one rate: the sum of its
         * sources in the list's order, then times -V (:829-835) */
        const int s = pr.srcOfCell[I];
        if (s >= 0) {
            E3 q[3];
            sourceOf(pr.pvt, qi, pr.P->src_type[s], pr.P->src_a[s], pr.P->src_b[s], V, q);
            for (int s2 = pr.srcNext[s]; s2 >= 0; s2 = pr.srcNext[s2]) {
                E3 q2[3];
                sourceOf(pr.pvt, qi, pr.P->src_type[s2], pr.P->src_a[s2], pr.P->src_b[s2], V, q2);
                for (int e = 0; e < 3; ++e) q[e] = q[e] + q2[e];
            }
            for (int e = 0; e < 3; ++e) {
                E3 t = q[e] * (-V);
                r[e] += t.v;
                for (int pvi = 0; pvi < 3; ++pvi) D[3 * e + pvi] += t.d[pvi];
            }
        }
    }
    /* boundary terms after the cell loop, the faces with a condition in cell
     * and face order (tpfalinearizer.hh:849-865) */
    for (const Problem::Bc& bc : pr.bcs) {
        if (bc.type == 0) continue;
        IQ<E3> qi;
        pr.iq<E3>(pv, mean, bc.cell, qi);
        E3 f[3];
        boundaryFlux(pr.pvt, qi, bc, pr.G.area(bc.face), pr.G.zc(bc.cell), pr.G.g, f);
        double* r = res + 3 * bc.cell;
        double* D = jac + 9L * pr.pat.diagOff[bc.cell];
        for (int e = 0; e < 3; ++e) {
            r[e] += f[e].v;
            for (int pvi = 0; pvi < 3; ++pvi) D[3 * e + pvi] += f[e].d[pvi];
        }
    }
}

/* ------------------------------------------------------------------ */
/* Block helpers, bs = 2 or 3, row-major (Dune FieldMatrix). */
template <class S>
bool invertBlock(int bs, S* a)
{
    if (bs == 2) {
        S det = a[0] * a[3] - a[1] * a[2];
        if (!(det != S(0)) || !std::isfinite((double)det)) return false;
        S r = S(1) / det;
        S i00 = a[3] * r, i01 = -a[1] * r, i10 = -a[2] * r, i11 = a[0] * r;
        a[0] = i00; a[1] = i01; a[2] = i10; a[3] = i11;
        return true;
    }
    S c00 = a[4] * a[8] - a[5] * a[7];
    S c01 = a[5] * a[6] - a[3] * a[8];
    S c02 = a[3] * a[7] - a[4] * a[6];
    S det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (!(det != S(0)) || !std::isfinite((double)det)) return false;
    S r = S(1) / det;
    S c10 = a[2] * a[7] - a[1] * a[8];
    S c11 = a[0] * a[8] - a[2] * a[6];
    S c12 = a[1] * a[6] - a[0] * a[7];
    S c20 = a[1] * a[5] - a[2] * a[4];
    S c21 = a[2] * a[3] - a[0] * a[5];
    S c22 = a[0] * a[4] - a[1] * a[3];
    a[0] = c00 * r; a[1] = c10 * r; a[2] = c20 * r;
    a[3] = c01 * r; a[4] = c11 * r; a[5] = c21 * r;
    a[6] = c02 * r; a[7] = c12 * r; a[8] = c22 * r;
    return true;
}
/* C = A * B */
template <class S>
void matmul(int bs, const S* A, const S* B, S* C)
{
    for (int i = 0; i < bs; ++i)
        for (int j = 0; j < bs; ++j) {
            S s = A[i * bs] * B[j];
            for (int k = 1; k < bs; ++k) s += A[i * bs + k] * B[k * bs + j];
            C[i * bs + j] = s;
        }
}
/* y += sign * A x  (in the order y_i = y_i + sign * sum_k A_ik x_k) */
template <class S>
void matvecAcc(int bs, const S* A, const S* x, S* y, S sign)
{
    for (int i = 0; i < bs; ++i) {
        S s = A[i * bs] * x[0];
        for (int k = 1; k < bs; ++k) s += A[i * bs + k] * x[k];
        y[i] += sign * s;
    }
}

/* Block ILU(0): dune-istl bilu0_decomposition as wrapped by
 * PreconditionerWrapperILU (istlpreconditionerwrappers.hh:151-190); diagonal
 * blocks are stored inverted. */
template <class S>
int ilu0Factor(int n, int bs, const int32_t* rp, const int32_t* ci, S* v)
{
    const int bb = bs * bs;
    std::vector<S> tmp(bb);
    for (int i = 0; i < n; ++i) {
        int ij = rp[i];
        for (; ij < rp[i + 1] && ci[ij] < i; ++ij) {
            const int j = ci[ij];
            /* A_jj (inverted) */
            int jj = rp[j];
            while (ci[jj] != j) ++jj;
            matmul(bs, v + (long)bb * ij, v + (long)bb * jj, tmp.data());
            std::copy(tmp.begin(), tmp.end(), v + (long)bb * ij);
            int ik = ij + 1, jk = jj + 1;
            while (ik < rp[i + 1] && jk < rp[j + 1]) {
                if (ci[ik] == ci[jk]) {
                    matmul(bs, v + (long)bb * ij, v + (long)bb * jk, tmp.data());
                    for (int t = 0; t < bb; ++t) v[(long)bb * ik + t] -= tmp[t];
                    ++ik; ++jk;
                } else if (ci[ik] < ci[jk]) ++ik;
                else ++jk;
            }
        }
        if (ij >= rp[i + 1] || ci[ij] != i) return -1;
        if (!invertBlock(bs, v + (long)bb * ij)) return 2;
    }
    return 0;
}

/* bilu_backsolve: forward with unit lower, backward with inverted diagonal,
 * then v *= relax (SeqILU::apply). */
template <class S>
void ilu0Apply(int n, int bs, const int32_t* rp, const int32_t* ci, const S* lu, const S* d,
               S* v, S relax)
{
    const int bb = bs * bs;
    for (int i = 0; i < n; ++i) {
        S rhs[3];
        for (int e = 0; e < bs; ++e) rhs[e] = d[bs * i + e];
        for (int k = rp[i]; k < rp[i + 1] && ci[k] < i; ++k)
            matvecAcc(bs, lu + (long)bb * k, v + (long)bs * ci[k], rhs, S(-1));
        for (int e = 0; e < bs; ++e) v[bs * i + e] = rhs[e];
    }
    for (int i = n - 1; i >= 0; --i) {
        S rhs[3];
        for (int e = 0; e < bs; ++e) rhs[e] = v[bs * i + e];
        int diag = -1;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            if (ci[k] == i) diag = k;
            else if (ci[k] > i) matvecAcc(bs, lu + (long)bb * k, v + (long)bs * ci[k], rhs, S(-1));
        }
        S out[3] = {0, 0, 0};
        matvecAcc(bs, lu + (long)bb * diag, rhs, out, S(1));
        for (int e = 0; e < bs; ++e) v[bs * i + e] = out[e];
    }
    if (relax != S(1))
        for (long k = 0; k < (long)n * bs; ++k) v[k] *= relax;
}

template <class S>
void spmv(int n, int bs, const int32_t* rp, const int32_t* ci, const S* a, const S* x, S* y)
{
    const int bb = bs * bs;
    for (int i = 0; i < n; ++i) {
        S acc[3] = {0, 0, 0};
        for (int k = rp[i]; k < rp[i + 1]; ++k) matvecAcc(bs, a + (long)bb * k, x + (long)bs * ci[k], acc, S(1));
        for (int e = 0; e < bs; ++e) y[bs * i + e] = acc[e];
    }
}

/* Exact sum of doubles, rounded once to the nearest double (ties to even).
 * The reference's scalar product is a sequential loop
 * (overlappingscalarproduct.hh:58-70); its value depends on the order of the
 * additions, so the restatement uses the exact sum, which that loop
 * approximates and which any summation order reproduces.  Positive and
 * negative products accumulate separately as unsigned magnitudes in 64-bit
 * digits of 2^-1074 units (34 digits hold any finite double with headroom);
 * the signed difference is rounded at the end.  A zero sum is +0; a NaN
 * product, or +inf with -inf, gives NaN; otherwise an infinite product gives
 * that infinity. */
class ExactSum {
public:
    void add(double v)
    {
        if (std::isnan(v)) { nan_ = true; return; }
        if (std::isinf(v)) { (v > 0 ? pinf_ : ninf_) = true; return; }
        if (v == 0.0) return;
        int ex = 0;
        const double f = std::frexp(std::fabs(v), &ex);         /* |v| = f 2^ex, f in [0.5, 1) */
        uint64_t m = (uint64_t)std::ldexp(f, 53);                /* integer: |v| = m 2^(ex - 53) */
        int p = ex - 53 + 1074;                                  /* position of m's unit bit */
        while (p < 0) { m >>= 1; ++p; }                          /* subnormals: exact shifts */
        uint64_t* acc = v > 0 ? pos_ : neg_;
        const int d = p / 64, s = p % 64;
        const unsigned __int128 w = (unsigned __int128)m << s;
        addAt(acc, d, (uint64_t)w);
        addAt(acc, d + 1, (uint64_t)(w >> 64));
    }
    double value() const
    {
        if (nan_ || (pinf_ && ninf_)) return std::numeric_limits<double>::quiet_NaN();
        if (pinf_) return std::numeric_limits<double>::infinity();
        if (ninf_) return -std::numeric_limits<double>::infinity();
        /* |pos - neg| and its sign */
        int c = 0;
        for (int k = kDigits - 1; k >= 0 && c == 0; --k)
            c = pos_[k] > neg_[k] ? 1 : (pos_[k] < neg_[k] ? -1 : 0);
        if (c == 0) return 0.0;
        const uint64_t* a = c > 0 ? pos_ : neg_;
        const uint64_t* b = c > 0 ? neg_ : pos_;
        uint64_t dif[kDigits];
        uint64_t borrow = 0;
        for (int k = 0; k < kDigits; ++k) {
            const unsigned __int128 t = (unsigned __int128)a[k] - b[k] - borrow;
            dif[k] = (uint64_t)t;
            borrow = (uint64_t)(t >> 127);
        }
        int top = kDigits - 1;
        while (dif[top] == 0) --top;
        const int topBit = 64 * top + 63 - __builtin_clzll(dif[top]);   /* in units of 2^-1074 */
        double r;
        if (topBit < 53) {
            r = std::ldexp((double)dif[0], -1074);                      /* exact: below 2^53 units */
        } else {
            /* the 53 leading bits, the next bit, and whether anything is below */
            const int low = topBit - 52;                                 /* unit bit of the mantissa */
            auto bit = [&](int i) -> uint64_t { return (dif[i / 64] >> (i % 64)) & 1u; };
            uint64_t mant = 0;
            for (int i = topBit; i >= low; --i) mant = (mant << 1) | bit(i);
            const bool half = bit(low - 1) != 0;
            bool below = false;
            for (int i = low - 2; i >= 0 && !below; --i) below = bit(i) != 0;
            if (half && (below || (mant & 1u))) ++mant;
            r = std::ldexp((double)mant, low - 1074);                   /* inf when too large */
        }
        return c > 0 ? r : -r;
    }

private:
    static constexpr int kDigits = 34;
    static void addAt(uint64_t* acc, int d, uint64_t v)
    {
        for (; v && d < kDigits; ++d) {
            const uint64_t o = acc[d];
            acc[d] = o + v;
            v = acc[d] < o ? 1 : 0;
        }
    }
    uint64_t pos_[kDigits] = {0}, neg_[kDigits] = {0};
    bool nan_ = false, pinf_ = false, ninf_ = false;
};

/* Scalar product.  Double: the exact sum of the rounded products (ExactSum).
 * Float (LinearSolverScalar = float): the reference's sequential loop in float. */
template <class S>
S dot(const std::vector<S>& x, const std::vector<S>& y)
{
    if constexpr (sizeof(S) == 8) {
        ExactSum e;
        for (size_t i = 0; i < x.size(); ++i) e.add(x[i] * y[i]);
        return e.value();
    } else {
        S s = 0;
        for (size_t i = 0; i < x.size(); ++i) s += x[i] * y[i];
        return s;
    }
}

/* CombinedCriterion (combinedcriterion.hh:54-236). */
template <class S>
struct Criterion {
    S tol, absTol, maxErr, res, init, last; bool stag;
    void updateErrors(const std::vector<S>& change, const std::vector<S>& r)
    {
        last = res; res = 0; stag = true;
        for (size_t i = 0; i < r.size(); ++i) {
            res = std::max<S>(res, std::abs(r[i]));
            if (stag && change[i] != S(0)) stag = false;
        }
    }
    void setInitial(const std::vector<S>& x, const std::vector<S>& r)
    {
        updateErrors(x, r);
        stag = false;
        res = std::max<S>(res, std::numeric_limits<S>::min() * S(1e10));
        init = res; last = res;
    }
    bool converged() const { return res / std::max<S>(S(1e-20), init) <= tol || res <= absTol; }
    bool failed() const { return !converged() && (stag || res > maxErr); }
};

struct Amg;
struct Precond {
    int kind, bs, n; double relax; int order;
    const int32_t *rp, *ci;
    int nx = 0, ny = 0, nz = 0;     /* grid of the two-colour tile ordering (kind 2) */
    const Amg* amg = nullptr;       /* kind 3 */
};

/* ------------------------------------------------------------------ */
/* Block ILU(0) in a two-colour tile ordering: the preconditioner of the
 * product's FP32-inner and ensemble forms (the reference has only the
 * natural order, PreconditionerWrapperILU, istlpreconditionerwrappers.hh:
 * 151-190; this restates the product's variant so that its kernels can be
 * checked bit for bit).  Tiles of 4 x 2 x 8 cells, coloured (tx + ty + tz)
 * mod 2; the order is: the cells of colour-0 tiles, then those of colour-1
 * tiles, natural inside a tile.  On the 7-point pattern ILU(0) in any order
 * only changes the diagonal blocks:
 *     D_I = A_II - sum_{K lower} (A_IK D_K^-1) A_KI,
 * M = (D + L_A) D^-1 (D + U_A).  Operation order per cell: the neighbours in
 * other tiles in column order, then those of its own tile in column order. */
struct TileColoring {
    int nx, ny, nz;
    int TX = 4, TY = 2, TZ = 8;       /* tile shape (1 x 1 x 1: red-black cells) */
    int ntx, nty;
    TileColoring(int x, int y, int z, int tx = 4, int ty = 2, int tz = 8)
        : nx(x), ny(y), nz(z), TX(tx), TY(ty), TZ(tz), ntx((x + tx - 1) / tx), nty((y + ty - 1) / ty) {}
    void ijk(long I, int& i, int& j, int& k) const
    {
        i = (int)(I % nx); j = (int)((I / nx) % ny); k = (int)(I / ((long)nx * ny));
    }
    long tile(long I) const
    {
        int i, j, k; ijk(I, i, j, k);
        return i / TX + (long)ntx * (j / TY + (long)nty * (k / TZ));
    }
    int color(long I) const
    {
        int i, j, k; ijk(I, i, j, k);
        return (i / TX + j / TY + k / TZ) & 1;
    }
    /* J comes before I in the ordering */
    bool before(long J, long I) const
    {
        if (tile(J) == tile(I)) return J < I;
        return color(J) < color(I);
    }
    /* the cells in the ordering */
    std::vector<long> order() const
    {
        std::vector<long> o;
        const long N = (long)nx * ny * nz;
        for (int c = 0; c < 2; ++c)
            for (long t = 0, nt = (long)ntx * nty * ((nz + TZ - 1) / TZ); t < nt; ++t) {
                const int tx = (int)(t % ntx), ty = (int)((t / ntx) % nty), tz = (int)(t / ((long)ntx * nty));
                if (((tx + ty + tz) & 1) != c) continue;
                for (int k = tz * TZ; k < std::min(nz, tz * TZ + TZ); ++k)
                    for (int j = ty * TY; j < std::min(ny, ty * TY + TY); ++j)
                        for (int i = tx * TX; i < std::min(nx, tx * TX + TX); ++i)
                            o.push_back(i + (long)nx * (j + (long)ny * k));
            }
        (void)N;
        return o;
    }
    /* neighbours of I in column order, split into other-tile and own-tile ones */
    void neighbours(long I, std::vector<long>& other, std::vector<long>& own) const
    {
        int i, j, k; ijk(I, i, j, k);
        const long nxy = (long)nx * ny;
        long cand[6]; int nc = 0;
        if (k > 0) cand[nc++] = I - nxy;
        if (j > 0) cand[nc++] = I - nx;
        if (i > 0) cand[nc++] = I - 1;
        if (i + 1 < nx) cand[nc++] = I + 1;
        if (j + 1 < ny) cand[nc++] = I + nx;
        if (k + 1 < nz) cand[nc++] = I + nxy;
        other.clear(); own.clear();
        for (int q = 0; q < nc; ++q) (tile(cand[q]) == tile(I) ? own : other).push_back(cand[q]);
    }
};

long blockAt(const int32_t* rp, const int32_t* ci, long row, long col)
{
    for (long k = rp[row]; k < rp[row + 1]; ++k)
        if (ci[k] == col) return k;
    return -1;
}

/* pivots D_I^-1 [n][9]; 2 for a singular pivot */
int coloredIluFactor(const TileColoring& tc, const int32_t* rp, const int32_t* ci, const double* A, double* dinv)
{
    std::vector<long> other, own;
    for (long I : tc.order()) {
        double D[9], P[9], T[9];
        std::copy(A + 9 * blockAt(rp, ci, I, I), A + 9 * blockAt(rp, ci, I, I) + 9, D);
        tc.neighbours(I, other, own);
        auto eliminate = [&](long K) {
            matmul(3, A + 9 * blockAt(rp, ci, I, K), dinv + 9 * K, P);
            matmul(3, P, A + 9 * blockAt(rp, ci, K, I), T);
            for (int q = 0; q < 9; ++q) D[q] -= T[q];
        };
        for (long K : other) if (tc.before(K, I)) eliminate(K);
        for (long K : own) if (tc.before(K, I)) eliminate(K);
        if (!invertBlock(3, D)) return 2;
        std::copy(D, D + 9, dinv + 9 * I);
    }
    return 0;
}

/* v = M^-1 d, then v *= relax */
void coloredIluApply(const TileColoring& tc, const int32_t* rp, const int32_t* ci, const double* A,
                     const double* dinv, const double* d, double* v, double relax)
{
    std::vector<long> other, own;
    const std::vector<long> ord = tc.order();
    for (long I : ord) {
        double acc[3] = {d[3 * I], d[3 * I + 1], d[3 * I + 2]};
        tc.neighbours(I, other, own);
        for (long K : other) if (tc.before(K, I)) matvecAcc(3, A + 9 * blockAt(rp, ci, I, K), v + 3 * K, acc, -1.0);
        for (long K : own) if (tc.before(K, I)) matvecAcc(3, A + 9 * blockAt(rp, ci, I, K), v + 3 * K, acc, -1.0);
        double w[3] = {0, 0, 0};
        matvecAcc(3, dinv + 9 * I, acc, w, 1.0);
        for (int e = 0; e < 3; ++e) v[3 * I + e] = w[e];
    }
    for (auto it = ord.rbegin(); it != ord.rend(); ++it) {
        const long I = *it;
        double acc[3] = {0, 0, 0};
        tc.neighbours(I, other, own);
        for (long J : other) if (tc.before(I, J)) matvecAcc(3, A + 9 * blockAt(rp, ci, I, J), v + 3 * J, acc, -1.0);
        for (long J : own) if (tc.before(I, J)) matvecAcc(3, A + 9 * blockAt(rp, ci, I, J), v + 3 * J, acc, -1.0);
        double o[3] = {0, 0, 0};
        matvecAcc(3, dinv + 9 * I, acc, o, 1.0);
        for (int e = 0; e < 3; ++e) v[3 * I + e] = v[3 * I + e] + o[e];
    }
    if (relax != 1.0)
        for (long k = 0; k < 3L * tc.nx * tc.ny * tc.nz; ++k) v[k] *= relax;
}

/* ------------------------------------------------------------------ */
/* Aggregation AMG of a 3x3 system on the 7-point grid: the product's
 * preconditioner in place of the reference's ParallelAmgBackend
 * (parallelamgbackend.hh:85-316: BiCGStab preconditioned by dune-istl's
 * aggregation AMG, SOR smoother with one iteration and relaxation 1 (:272-276),
 * aggregates of about 3 cells per direction (:294-295), coarsening until
 * AmgCoarsenTarget (:72, :290)).  dune-istl is not part of the reference, so
 * its aggregation (strength of connection by the blocks' Frobenius norm), its
 * cycle and its coarse solver are not restated; this is the product's own
 * algorithm, which the product's kernels follow bit for bit:
 *   - box aggregates of ax x ay x az cells (cell (i,j,k) -> (i/ax, j/ay,
 *     k/az)), levels until one has at most `target` cells;
 *   - Galerkin coarse operators P^T A P with piecewise-constant P: a coarse
 *     block is the sum of the fine blocks between the two aggregates (again a
 *     7-point pattern), summed over the fine cells of the aggregate in natural
 *     order and, per cell, over its slots in order -z, -y, -x, I, +x, +y, +z;
 *   - the V-cycle on level l for rhs b: x = 0; `pre` forward block
 *     Gauss-Seidel sweeps (the first from x = 0); r = b - A x on the even
 *     cells (the odd ones, swept last, have r = 0 but for rounding); the
 *     coarse rhs is the sum of r over each aggregate's even cells; the
 *     coarse solution xc; x += damp * P xc; `post` backward sweeps;
 *   - Gauss-Seidel in red-black order (the cells with (i + j + k) even
 *     first, each colour in natural order; backward: the reverse):
 *     x_I = A_II^-1 (b_I - sum_{K != I} A_IK x_K) with the current x, the
 *     neighbours in column order (all of them of the other colour);
 *   - the coarsest level: x = M^-1 b with M^-1 by in-place Gauss-Jordan with
 *     partial pivoting (the first largest |entry| of the column), rows
 *     summed in column order. */
struct AmgLevel {
    int nx, ny, nz;
    long N;
    int ax = 1, ay = 1, az = 1;        /* aggregation to the next level */
    std::vector<double> A;             /* [7][N][9], slots -z, -y, -x, I, +x, +y, +z */
    std::vector<double> dinv;          /* [N][9] inverted diagonal blocks */
};
struct Amg {
    std::vector<AmgLevel> lv;
    std::vector<double> dense;         /* the coarsest level's inverse, [n][n] row-major */
    int pre = 2, post = 2;
    double damp = 1.6;
};

/* neighbour of I through slot s (0..6), or -1 */
long slotNeighbour(int nx, int ny, int nz, long I, int s)
{
    const int i = (int)(I % nx), j = (int)((I / nx) % ny), k = (int)(I / ((long)nx * ny));
    const long nxy = (long)nx * ny;
    switch (s) {
    case 0: return k > 0 ? I - nxy : -1;
    case 1: return j > 0 ? I - nx : -1;
    case 2: return i > 0 ? I - 1 : -1;
    case 3: return I;
    case 4: return i + 1 < nx ? I + 1 : -1;
    case 5: return j + 1 < ny ? I + nx : -1;
    default: return k + 1 < nz ? I + nxy : -1;
    }
}

/* the sorted 7-point BCRS rows in the structured layout (absent blocks zero) */
std::vector<double> diaFromBcrs(int nx, int ny, int nz, const int32_t* rp, const int32_t* ci, const double* J)
{
    const long N = (long)nx * ny * nz;
    std::vector<double> A(63 * N, 0.0);
    for (long I = 0; I < N; ++I)
        for (int s = 0; s < 7; ++s) {
            const long K = slotNeighbour(nx, ny, nz, I, s);
            if (K < 0) continue;
            const long b = blockAt(rp, ci, I, K);
            if (b >= 0) std::copy(J + 9 * b, J + 9 * b + 9, A.begin() + 9 * (s * N + I));
        }
    return A;
}

/* level dims: aggregates of agg[] cells until a level has at most target
 * cells (or no direction can be coarsened, or 16 levels) */
std::vector<std::array<int, 3>> amgDims(int nx, int ny, int nz, const int agg[3], int target)
{
    std::vector<std::array<int, 3>> d{{nx, ny, nz}};
    while ((long)d.back()[0] * d.back()[1] * d.back()[2] > target && d.size() < 16) {
        const auto& f = d.back();
        std::array<int, 3> c{(f[0] + agg[0] - 1) / agg[0], (f[1] + agg[1] - 1) / agg[1], (f[2] + agg[2] - 1) / agg[2]};
        if ((long)c[0] * c[1] * c[2] == (long)f[0] * f[1] * f[2]) break;
        d.push_back(c);
    }
    return d;
}

/* coarse A from the fine level f (aggregation f.ax x f.ay x f.az) */
void amgGalerkin(const AmgLevel& f, AmgLevel& c)
{
    const long nxy = (long)f.nx * f.ny;
    for (long a = 0; a < c.N; ++a) {
        const int ci = (int)(a % c.nx), cj = (int)((a / c.nx) % c.ny), ck = (int)(a / ((long)c.nx * c.ny));
        for (int e = 0; e < 9; ++e) {
            double acc[7] = {0, 0, 0, 0, 0, 0, 0};
            for (int k = ck * f.az; k < std::min(f.nz, ck * f.az + f.az); ++k)
                for (int j = cj * f.ay; j < std::min(f.ny, cj * f.ay + f.ay); ++j)
                    for (int i = ci * f.ax; i < std::min(f.nx, ci * f.ax + f.ax); ++i) {
                        const long I = i + (long)f.nx * j + nxy * k;
                        for (int s = 0; s < 7; ++s) {
                            const long J = slotNeighbour(f.nx, f.ny, f.nz, I, s);
                            if (J < 0) continue;
                            const int ji = (int)(J % f.nx), jj = (int)((J / f.nx) % f.ny), jk = (int)(J / nxy);
                            const bool same = ji / f.ax == ci && jj / f.ay == cj && jk / f.az == ck;
                            acc[same ? 3 : s] += f.A[9 * (s * f.N + I) + e];
                        }
                    }
            for (int t = 0; t < 7; ++t) c.A[9 * (t * c.N + a) + e] = acc[t];
        }
    }
}

/* one Gauss-Seidel sweep, in place (zero: x = 0 before the sweep) */
void amgGsSweep(const AmgLevel& L, const double* b, double* x, bool forward, bool zero)
{
    const TileColoring tc(L.nx, L.ny, L.nz, 1, 1, 1);
    std::vector<long> ord = tc.order();
    if (!forward) std::reverse(ord.begin(), ord.end());
    std::vector<char> done(L.N, 0);
    std::vector<long> other, own;
    auto blockOf = [&](long I, long K) {
        for (int s = 0; s < 7; ++s)
            if (s != 3 && slotNeighbour(L.nx, L.ny, L.nz, I, s) == K) return L.A.data() + 9 * (s * L.N + I);
        return (const double*)nullptr;
    };
    for (long I : ord) {
        double acc[3] = {b[3 * I], b[3 * I + 1], b[3 * I + 2]};
        tc.neighbours(I, other, own);
        for (long K : other) if (done[K] || !zero) matvecAcc(3, blockOf(I, K), x + 3 * K, acc, -1.0);
        for (long K : own) if (done[K] || !zero) matvecAcc(3, blockOf(I, K), x + 3 * K, acc, -1.0);
        double w[3] = {0, 0, 0};
        matvecAcc(3, L.dinv.data() + 9 * I, acc, w, 1.0);
        for (int e = 0; e < 3; ++e) x[3 * I + e] = w[e];
        done[I] = 1;
    }
}

/* r = b - A x, per row: the slots in order */
/* r = b - A x of the even cells (i + j + k even); the odd cells, swept last,
 * solve their rows with their neighbours' final values, so their residual
 * is zero but for rounding and is neither formed nor restricted */
bool evenCell(const AmgLevel& L, long I)
{
    return ((I % L.nx + (I / L.nx) % L.ny + I / ((long)L.nx * L.ny)) & 1) == 0;
}

void amgResidual(const AmgLevel& L, const double* b, const double* x, double* r)
{
    for (long I = 0; I < L.N; ++I) {
        if (!evenCell(L, I)) continue;
        double acc[3] = {b[3 * I], b[3 * I + 1], b[3 * I + 2]};
        for (int s = 0; s < 7; ++s) {
            const long J = slotNeighbour(L.nx, L.ny, L.nz, I, s);
            if (J >= 0) matvecAcc(3, L.A.data() + 9 * (s * L.N + I), x + 3 * J, acc, -1.0);
        }
        for (int e = 0; e < 3; ++e) r[3 * I + e] = acc[e];
    }
}

/* coarse rhs: the sum of r over each aggregate's even cells, in natural order */
void amgRestrict(const AmgLevel& f, const AmgLevel& c, const double* r, double* bc)
{
    for (long a = 0; a < c.N; ++a) {
        const int ci = (int)(a % c.nx), cj = (int)((a / c.nx) % c.ny), ck = (int)(a / ((long)c.nx * c.ny));
        for (int e = 0; e < 3; ++e) {
            double acc = 0.0;
            for (int k = ck * f.az; k < std::min(f.nz, ck * f.az + f.az); ++k)
                for (int j = cj * f.ay; j < std::min(f.ny, cj * f.ay + f.ay); ++j)
                    for (int i = ci * f.ax; i < std::min(f.nx, ci * f.ax + f.ax); ++i)
                        if (((i + j + k) & 1) == 0) acc += r[3 * (i + (long)f.nx * (j + (long)f.ny * k)) + e];
            bc[3 * a + e] = acc;
        }
    }
}

/* x += damp * P xc */
void amgProlong(const AmgLevel& f, const AmgLevel& c, const double* xc, double* x, double damp)
{
    for (long I = 0; I < f.N; ++I) {
        const int i = (int)(I % f.nx), j = (int)((I / f.nx) % f.ny), k = (int)(I / ((long)f.nx * f.ny));
        const long a = i / f.ax + (long)c.nx * (j / f.ay + (long)c.ny * (k / f.az));
        for (int e = 0; e < 3; ++e) x[3 * I + e] = x[3 * I + e] + damp * xc[3 * a + e];
    }
}

/* in-place Gauss-Jordan inverse with partial pivoting; 2 when singular or
 * not finite */
int denseInverse(int n, double* M)
{
    for (long q = 0; q < (long)n * n; ++q)
        if (!std::isfinite(M[q])) return 2;
    std::vector<int> piv(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::abs(M[(long)k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (std::abs(M[(long)i * n + k]) > best) { best = std::abs(M[(long)i * n + k]); p = i; }
        if (!(best != 0.0)) return 2;
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(M[(long)k * n + j], M[(long)p * n + j]);
        const double r = 1.0 / M[(long)k * n + k];
        M[(long)k * n + k] = 1.0;
        for (int j = 0; j < n; ++j) M[(long)k * n + j] *= r;
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            const double f = M[(long)i * n + k];
            M[(long)i * n + k] = 0.0;
            for (int j = 0; j < n; ++j) M[(long)i * n + j] -= M[(long)k * n + j] * f;
        }
    }
    for (int k = n - 1; k >= 0; --k)
        if (piv[k] != k)
            for (int i = 0; i < n; ++i) std::swap(M[(long)i * n + k], M[(long)i * n + piv[k]]);
    return 0;
}

/* hierarchy from the fine matrix in the structured layout; 2 for a singular
 * diagonal block or coarsest matrix */
int amgSetup(Amg& amg, int nx, int ny, int nz, std::vector<double> A0, const int agg[3], int target)
{
    const auto dims = amgDims(nx, ny, nz, agg, target);
    amg.lv.clear();
    for (size_t l = 0; l < dims.size(); ++l) {
        AmgLevel L;
        L.nx = dims[l][0]; L.ny = dims[l][1]; L.nz = dims[l][2];
        L.N = (long)L.nx * L.ny * L.nz;
        if (l + 1 < dims.size()) { L.ax = agg[0]; L.ay = agg[1]; L.az = agg[2]; }
        if (l == 0) L.A = std::move(A0);
        else { L.A.assign(63 * L.N, 0.0); amgGalerkin(amg.lv.back(), L); }
        L.dinv.resize(9 * L.N);
        for (long I = 0; I < L.N; ++I) {
            std::copy(L.A.begin() + 9 * (3 * L.N + I), L.A.begin() + 9 * (3 * L.N + I) + 9, L.dinv.begin() + 9 * I);
            if (l + 1 < dims.size() && !invertBlock(3, L.dinv.data() + 9 * I)) return 2;
        }
        amg.lv.push_back(std::move(L));
    }
    const AmgLevel& C = amg.lv.back();
    const int n = (int)(3 * C.N);
    amg.dense.assign((long)n * n, 0.0);
    for (long I = 0; I < C.N; ++I)
        for (int s = 0; s < 7; ++s) {
            const long J = slotNeighbour(C.nx, C.ny, C.nz, I, s);
            if (J < 0) continue;
            for (int e = 0; e < 3; ++e)
                for (int f = 0; f < 3; ++f) amg.dense[(3 * I + e) * (long)n + 3 * J + f] = C.A[9 * (s * C.N + I) + 3 * e + f];
        }
    return denseInverse(n, amg.dense.data());
}

void amgCycle(const Amg& amg, size_t l, const double* b, double* x)
{
    const AmgLevel& L = amg.lv[l];
    if (l + 1 == amg.lv.size()) {
        const long n = 3 * L.N;
        for (long i = 0; i < n; ++i) {
            double s = amg.dense[i * n] * b[0];
            for (long j = 1; j < n; ++j) s += amg.dense[i * n + j] * b[j];
            x[i] = s;
        }
        return;
    }
    const AmgLevel& C = amg.lv[l + 1];
    for (int k = 0; k < amg.pre; ++k) amgGsSweep(L, b, x, true, k == 0);
    std::vector<double> r(3 * L.N), bc(3 * C.N), xc(3 * C.N);
    amgResidual(L, b, x, r.data());
    amgRestrict(L, C, r.data(), bc.data());
    amgCycle(amg, l + 1, bc.data(), xc.data());
    amgProlong(L, C, xc.data(), x, amg.damp);
    for (int k = 0; k < amg.post; ++k) amgGsSweep(L, b, x, false, false);
}

/* Block Jacobi, Dune::SeqJac(A, order, w) as PreconditionerWrapperJacobi
 * builds it (istlpreconditionerwrappers.hh:84-89): `order` sweeps
 * x <- (1 - w) x + w D^{-1} (d - (L + U) x) starting from the output vector's
 * current content (BiCGStab passes y, which holds the previous t = A z, and
 * z; both start at zero).  Order 0 leaves the output as it is (the reference
 * default, linalgparameters.hh:74).  M holds the inverted diagonal blocks.
 * The relaxed update's form, (1 - w) x + w x_J, is dune's to decide ("parity
 * unpinned" for w != 1; for w = 1 it is the plain Jacobi step). */
template <class S>
void jacobiSweeps(const Precond& pc, const S* A, const S* M, const std::vector<S>& in, std::vector<S>& out)
{
    const int bs = pc.bs, bb = bs * bs, n = pc.n;
    const int32_t *rp = pc.rp, *ci = pc.ci;
    std::vector<S> prev;
    for (int sweep = 0; sweep < pc.order; ++sweep) {
        prev = out;
        for (int i = 0; i < n; ++i) {
            S rhs[3];
            for (int e = 0; e < bs; ++e) rhs[e] = in[bs * i + e];
            for (int k = rp[i]; k < rp[i + 1]; ++k)
                if (ci[k] != i) matvecAcc(bs, A + (long)bb * k, prev.data() + (long)bs * ci[k], rhs, S(-1));
            S o[3] = {0, 0, 0};
            matvecAcc(bs, M + (long)bb * i, rhs, o, S(1));
            for (int e = 0; e < bs; ++e)
                out[bs * i + e] = (S(1) - (S)pc.relax) * prev[bs * i + e] + (S)pc.relax * o[e];
        }
    }
}

/* BiCGStabSolver::apply (bicgstabsolver.hh:120-330) with the aliasing of the
 * reference (h == x, s == r, t == y).  Returns 0 converged, 1 not converged,
 * 2 breakdown (NumericalProblem). */
template <class S>
int bicgstab(int n, int bs, const int32_t* rp, const int32_t* ci, const S* A, const S* M,
             const std::vector<S>& b, std::vector<S>& x, const Precond& pc, S tol, S absTol,
             S maxErr, int maxIter, int* itersOut)
{
    const long m = (long)n * bs;
    const S breakdownEps = std::numeric_limits<S>::min() * S(1e10);
    auto precApply = [&](std::vector<S>& out, const std::vector<S>& in) {
        if (pc.kind == 0) ilu0Apply<S>(n, bs, rp, ci, M, in.data(), out.data(), (S)pc.relax);
        else if (pc.kind == 1) jacobiSweeps<S>(pc, A, M, in, out);
        else if constexpr (sizeof(S) == 8) {
            if (pc.kind == 2)
                coloredIluApply(TileColoring(pc.nx, pc.ny, pc.nz), rp, ci, A, M, in.data(), out.data(), pc.relax);
            else
                amgCycle(*pc.amg, 0, in.data(), out.data());
        }
    };
    x.assign(m, S(0));
    std::vector<S> r = b;
    Criterion<S> crit{tol, absTol, maxErr, 0, 0, 0, false};
    crit.setInitial(x, r);
    int it = 0;
    *itersOut = 0;
    if (crit.converged()) return 0;
    const std::vector<S>& r0hat = b;
    S rho = 1, alpha = 1, omega = 1;
    std::vector<S> v(m, S(0)), p(m, S(0)), y(m, S(0)), z(m, S(0));
    for (; it < maxIter; ++it) {
        *itersOut = it;
        S rho_i = dot(r0hat, r);
        if (std::abs(rho) <= breakdownEps || std::abs(omega) <= breakdownEps) return 2;
        S beta = (rho_i / rho) * (alpha / omega);
        rho = rho_i;
        for (long i = 0; i < m; ++i) {
            S tmp = v[i];
            tmp *= omega;
            tmp -= p[i];
            tmp *= -beta;
            p[i] = r[i];
            p[i] += tmp;
        }
        precApply(y, p);
        spmv<S>(n, bs, rp, ci, A, y.data(), v.data());
        S denom = dot(r0hat, v);
        if (std::abs(denom) <= breakdownEps) return 2;
        alpha = rho_i / denom;
        if (std::abs(alpha) <= breakdownEps) return 2;
        for (long i = 0; i < m; ++i) {
            S tmp = y[i];
            tmp *= alpha;
            tmp += x[i];
            x[i] = tmp;
            tmp = v[i];
            tmp *= alpha;
            r[i] -= tmp;
        }
        crit.updateErrors(y, r);
        if (crit.converged()) { *itersOut = it; return 0; }
        if (crit.failed()) { *itersOut = it; return 1; }
        precApply(z, r);
        spmv<S>(n, bs, rp, ci, A, z.data(), y.data());
        denom = dot(y, y);
        if (std::abs(denom) <= breakdownEps) return 2;
        omega = dot(y, r) / denom;
        if (std::abs(omega) <= breakdownEps) return 2;
        for (long i = 0; i < m; ++i) x[i] += omega * z[i];
        crit.updateErrors(z, r);
        if (crit.converged()) { *itersOut = it; return 0; }
        if (crit.failed()) { *itersOut = it; return 1; }
        for (long i = 0; i < m; ++i) r[i] += -omega * y[i];
    }
    *itersOut = it;
    return 1;
}

bool amgParamsValid(const ob_params* p)
{
    return p->amg_coarsen_target >= 1 && p->amg_coarsen_target <= 32 && p->amg_pre_smooth >= 1 &&
           p->amg_post_smooth >= 0 && p->amg_agg[0] >= 1 && p->amg_agg[1] >= 1 && p->amg_agg[2] >= 1 &&
           std::isfinite(p->amg_prolong_damp);
}

/* ParallelBaseBackend::solve (parallelbasebackend.hh:252-284) +
 * ParallelBiCGStabSolverBackend::prepareSolver_ (parallelbicgstabbackend.hh:131-160):
 * copy (and convert to the solver scalar, overlappingbcrsmatrix.hh:220-262),
 * prepare the preconditioner, run, copy back. */
template <class S>
int backendSolve(int n, int bs, const int32_t* rp, const int32_t* ci, const double* J,
                 const double* b, double* xOut, const ob_params* prm, int* iters, const int* dims = nullptr)
{
    const int bb = bs * bs;
    const long nnzb = rp[n];
    std::vector<S> A(J, J + (long)bb * nnzb);
    std::vector<S> M;
    const bool colored = prm->precond == 0 && prm->ilu_order == 1;
    Amg amg;
    if (prm->precond == 2) {
        if constexpr (sizeof(S) == 8) {
            if (!dims || bs != 3 || (long)dims[0] * dims[1] * dims[2] != n || !amgParamsValid(prm)) return -1;
            amg.pre = prm->amg_pre_smooth; amg.post = prm->amg_post_smooth; amg.damp = prm->amg_prolong_damp;
            if (amgSetup(amg, dims[0], dims[1], dims[2], diaFromBcrs(dims[0], dims[1], dims[2], rp, ci, J), prm->amg_agg,
                         prm->amg_coarsen_target) != 0)
                return 2;
        } else {
            return -1;
        }
    } else if (colored) {
        if constexpr (sizeof(S) == 8) {
            if (!dims || bs != 3 || (long)dims[0] * dims[1] * dims[2] != n) return -1;
            M.resize((long)bb * n);
            if (coloredIluFactor(TileColoring(dims[0], dims[1], dims[2]), rp, ci, A.data(), M.data()) != 0) return 2;
        } else {
            return -1;
        }
    } else if (prm->precond == 0) {
        M = A;
        int rc = ilu0Factor<S>(n, bs, rp, ci, M.data());
        if (rc != 0) return 2;
    } else {
        M.resize((long)bb * n);
        for (int i = 0; i < n; ++i) {
            int k = rp[i];
            while (ci[k] != i) ++k;
            std::copy(A.begin() + (long)bb * k, A.begin() + (long)bb * (k + 1), M.begin() + (long)bb * i);
            if (!invertBlock(bs, M.data() + (long)bb * i)) return 2;
        }
    }
    std::vector<S> bv(b, b + (long)n * bs), x;
    Precond pc{colored ? 2 : (prm->precond == 2 ? 3 : prm->precond), bs, n, prm->prec_relax, prm->prec_order, rp, ci};
    if (dims) { pc.nx = dims[0]; pc.ny = dims[1]; pc.nz = dims[2]; }
    pc.amg = &amg;
    double absTol = prm->lin_abs_tol < 0 ? prm->newton_tol / 100.0 : prm->lin_abs_tol;
    int rc = bicgstab<S>(n, bs, rp, ci, A.data(), M.data(), bv, x, pc, (S)prm->lin_tol, (S)absTol,
                         (S)prm->lin_max_error, prm->lin_max_iter, iters);
    for (long k = 0; k < (long)n * bs; ++k) xOut[k] = (double)x[k];
    return rc;
}

/* ------------------------------------------------------------------ */
/* Newton update: NewtonMethod::update_ (newtonmethod.hh:665-710) ->
 * BlackOilNewtonMethod::updatePrimaryVariables_ (blackoilnewtonmethod.hh:226-427)
 * -> BlackOilPrimaryVariables::adaptPrimaryVariables (blackoilprimaryvariables.hh:549-909). */
bool adaptPrimary(const Pvt& pvt, double* x, uint8_t& mean, double eps)
{
    const double swMaximum = 1.0, waterOnlyThreshold = 1.0;
    double sw = x[0];
    double sg = (mean == MEAN_SG) ? x[2] : 0.0;
    if (sw >= waterOnlyThreshold) {
        x[0] = std::min(swMaximum, sw);
        x[2] = 0.0;
        bool changed = mean != MEAN_SG;
        mean = MEAN_SG;
        return changed;
    }
    bool changed = false;
    if (mean == MEAN_SG) {
        double s = 1.0 - sw;
        if (sg < -eps && s > 0.0) {
            const double po = x[1];
            mean = MEAN_RS;
            double rsSat = pvt.rsSat.eval(po);
            const double rsMax = std::numeric_limits<double>::max() / 2;
            x[2] = std::min(rsMax, rsSat);
            changed = true;
        }
    } else {
        const double po = x[1];
        double rsSat = pvt.rsSat.eval(po);
        const double rsMax = std::numeric_limits<double>::max() / 2;
        double rs = x[2];
        if (rs > std::min(rsMax, rsSat * (1.0 + eps))) {
            mean = MEAN_SG;
            x[2] = 0.0;
            changed = true;
        }
    }
    return changed;
}

void updatePrimary(const Pvt& pvt, const double* cur, uint8_t curMean, const double* upd,
                   double* next, uint8_t& nextMean, uint8_t& wasSw, const ob_params* prm)
{
    double dSw = 0.0, dSo = 0.0, dSg = 0.0;
    dSw = upd[0];
    dSo -= dSw;
    if (curMean == MEAN_SG) { dSg = upd[2]; dSo -= dSg; }
    double maxSatDelta = std::max(std::abs(dSg), std::abs(dSo));
    maxSatDelta = std::max(maxSatDelta, std::abs(dSw));
    double satAlpha = 1.0;
    if (maxSatDelta > prm->ds_max) satAlpha = prm->ds_max / maxSatDelta;
    for (int pvi = 0; pvi < 3; ++pvi) {
        double delta = upd[pvi];
        if (pvi == 1) {
            if (std::abs(delta) > prm->dp_max_rel * cur[pvi]) {
                double sgn = (delta > 0) - (delta < 0);
                delta = sgn * prm->dp_max_rel * cur[pvi];
            }
        } else if (pvi == 0) {
            delta *= satAlpha;
        } else {
            if (curMean == MEAN_SG) delta *= satAlpha;
            else if (delta > cur[2]) delta = cur[2];
        }
        next[pvi] = cur[pvi] - delta;
        if (pvi == 1) next[pvi] = std::clamp(next[pvi], -1e99, 1e99);
    }
    nextMean = curMean;
    const double eps = wasSw ? prm->privar_osc_thresh : 0.0;
    wasSw = adaptPrimary(pvt, next, nextMean, eps) ? 1 : 0;
}

} // namespace

/* ====================================================================== */
extern "C" {

void ob_default_params(ob_params* p)
{
    p->newton_tol = 1e-6; p->newton_max_iter = 20; p->newton_target_iter = 10;
    p->newton_max_error = 1e100; p->lin_tol = 1e-3; p->lin_abs_tol = -1.0;
    p->lin_max_error = 1e7; p->lin_max_iter = 1000; p->precond = 0; p->prec_relax = 1.0;
    p->prec_order = 0; p->max_dt_divisions = 10; p->dp_max_rel = 0.3; p->ds_max = 0.2;
    p->privar_osc_thresh = 1e-5; p->max_dt = std::numeric_limits<double>::infinity();
    p->min_dt = 0.0; p->solver_float = 0; p->verbose = 0; p->ilu_order = 0;
    p->amg_coarsen_target = 32; p->amg_pre_smooth = 2; p->amg_post_smooth = 2; p->amg_prolong_damp = 1.6;
    p->amg_agg[0] = p->amg_agg[1] = p->amg_agg[2] = 3;
}

int ob_pattern_size(const ob_problem* p, int64_t* nnzb, int64_t* nfaces2)
{
    try {
        Grid G(p);
        long f2 = 0;
        for (long I = 0; I < G.N; ++I)
            for (int d = 0; d < 6; ++d) if (G.nb(I, d) >= 0) ++f2;
        *nfaces2 = f2; *nnzb = G.N + f2;
        if (p->n_aux > 0) {
            Pattern P;
            buildPattern(G, P);
            *nnzb = P.rowPtr.back();
        }
        return 0;
    } catch (...) { return -1; }
}

int ob_boundary_flows(const ob_problem* p, const double* pv, const uint8_t* meaning, int32_t* cell, int32_t* face,
                      double* flows)
{
    try {
        Problem pr(p);
        const int n = (int)pr.bcs.size();
        for (int b = 0; b < n; ++b) {
            const Problem::Bc& bc = pr.bcs[b];
            if (cell) cell[b] = (int32_t)bc.cell;
            if (face) face[b] = bc.face;
            if (!flows) continue;
            for (int e = 0; e < 3; ++e) flows[3 * b + e] = 0.0;
            if (bc.type == 0) continue;
            IQ<E3> qi;
            pr.iq<E3>(pv, meaning, bc.cell, qi);
            E3 f[3];
            boundaryFlux(pr.pvt, qi, bc, pr.G.area(bc.face), pr.G.zc(bc.cell), pr.G.g, f);
            for (int e = 0; e < 3; ++e) flows[3 * b + e] = f[e].v;
        }
        return n;
    } catch (...) { return -1; }
}

int ob_add_auxiliary(const ob_problem* p, int32_t nb, const int32_t* brow, const int32_t* bcol,
                     const double* blocks, int32_t nr, const int32_t* rrow, const double* rvals, double* jac,
                     double* residual)
{
    try {
        Grid G(p);
        Pattern P;
        buildPattern(G, P);
        const long R = (long)P.rowPtr.size() - 1;
        for (int32_t k = 0; k < nb; ++k) {
            const long r = brow[k], c = bcol[k];
            if (r < 0 || r >= R) return -1;
            const int32_t* b = P.colIdx.data() + P.rowPtr[r];
            const int32_t* e = P.colIdx.data() + P.rowPtr[r + 1];
            const int32_t* it = std::lower_bound(b, e, (int32_t)c);
            if (it == e || *it != c) return -1;
            double* B = jac + 9L * (it - P.colIdx.data());
            for (int q = 0; q < 9; ++q) B[q] += blocks[9L * k + q];
        }
        for (int32_t k = 0; k < nr; ++k) {
            const long r = rrow[k];
            if (r < 0 || r >= R) return -1;
            for (int e = 0; e < 3; ++e) residual[3 * r + e] += rvals[3L * k + e];
        }
        return 0;
    } catch (...) { return -1; }
}

int ob_build_pattern(const ob_problem* p, int32_t* row_ptr, int32_t* col_idx, int32_t* diag_off,
                     int32_t* nb_ptr, int32_t* nb_cell, int32_t* nb_off, double* nb_trans,
                     double* nb_area, double* nb_dzg, int32_t* nb_dir)
{
    try {
        Grid G(p);
        Pattern P;
        buildPattern(G, P);
        std::copy(P.rowPtr.begin(), P.rowPtr.end(), row_ptr);
        std::copy(P.colIdx.begin(), P.colIdx.end(), col_idx);
        std::copy(P.diagOff.begin(), P.diagOff.end(), diag_off);
        std::copy(P.nbPtr.begin(), P.nbPtr.end(), nb_ptr);
        for (size_t k = 0; k < P.nb.size(); ++k) {
            nb_cell[k] = (int32_t)P.nb[k].J; nb_off[k] = P.nb[k].off;
            nb_trans[k] = P.nb[k].trans; nb_area[k] = P.nb[k].area;
            nb_dzg[k] = P.nb[k].dZg; nb_dir[k] = P.nb[k].dir;
        }
        return 0;
    } catch (...) { return -1; }
}

int ob_storage(const ob_problem* p, const double* pv, const uint8_t* meaning, double* storage)
{
    try {
        Grid G(p);
        Pvt pvt; pvt.init(p);
        const bool on = rockOn(p);
        const Rock rock = on ? makeRock(p) : Rock{};
        for (long I = 0; I < G.N; ++I) {
            IQ<double> q;
            computeIQ<double>(pvt, pv + 3 * I, meaning[I], p->poro[I], q);
            if (on) applyRock(rock, I, q);
            storageOf<double>(pvt, q, storage + 3 * I);
        }
        return 0;
    } catch (...) { return -1; }
}

int ob_intensive(const ob_problem* p, const double* pv3, int meaning, int cell, double* out)
{
    Pvt pvt; pvt.init(p);
    IQ<E3> q;
    computeIQ<E3>(pvt, pv3, meaning, p->poro[cell], q);
    if (rockOn(p)) applyRock(makeRock(p), cell, q);
    const E3* f[13] = {&q.p[0], &q.rho[W], &q.rho[O], &q.rho[G], &q.lam[W], &q.lam[O], &q.lam[G],
                       &q.invB[W], &q.invB[O], &q.invB[G], &q.Rs, &q.phi, &q.S[O]};
    for (int k = 0; k < 13; ++k) {
        out[4 * k] = f[k]->v;
        for (int i = 0; i < 3; ++i) out[4 * k + 1 + i] = f[k]->d[i];
    }
    return 0;
}

/* The problem's source terms as blocks: per source the residual contribution
 * (times -V) and its derivatives w.r.t. the cell's primaries, exactly what
 * linearize adds in its cell loop (tpfalinearizer.hh:827-841); with separate
 * sparse source terms they are added after the loop (:842-846). */
int ob_source_blocks(const ob_problem* p, const double* pv, const uint8_t* meaning, int32_t* cells,
                     double* residual, double* diag)
{
    try {
        Problem pr(p);
        for (int s = 0; s < p->n_src; ++s) {
            const long I = p->src_cell[s];
            const double V = pr.G.volume(I);
            IQ<E3> qi;
            pr.iq<E3>(pv, meaning, I, qi);
            E3 q[3];
            sourceOf(pr.pvt, qi, p->src_type[s], p->src_a[s], p->src_b[s], V, q);
            cells[s] = (int32_t)I;
            for (int e = 0; e < 3; ++e) {
                const E3 t = q[e] * (-V);
                residual[3 * s + e] = t.v;
                for (int pvi = 0; pvi < 3; ++pvi) diag[9 * s + 3 * e + pvi] = t.d[pvi];
            }
        }
        return p->n_src;
    } catch (...) { return -1; }
}

/* TpfaLinearizer::updateFlowsInfo (tpfalinearizer.hh:647-705): per cell and
 * neighbour in intersection order, the values of the face's component fluxes
 * (FLOWS) and the phases' reservoir-volume rates (FLORES), [entries][3]. */
int ob_flows(const ob_problem* p, const double* pv, const uint8_t* meaning, double* flows, double* flores)
{
    return ob_flows_ext(p, nullptr, nullptr, nullptr, pv, meaning, flows, flores);
}

int ob_flows_ext(const ob_problem* p, const ob_diffusion* dif, const ob_dispersion* dsp, const double* nv,
                 const double* pv, const uint8_t* meaning, double* flows, double* flores)
{
    try {
        Problem pr(p, dif, dsp);
        if (dsp && !nv) throw std::invalid_argument("dispersion needs the cells' velocities");
        for (long I = 0; I < pr.G.N; ++I) {
            IQ<E3> qi;
            pr.iq<E3>(pv, meaning, I, qi);
            for (int k = pr.pat.nbPtr[I]; k < pr.pat.nbPtr[I + 1]; ++k) {
                const NbInfo& nbi = pr.pat.nb[k];
                IQ<double> qj;
                pr.iq<double>(pv, meaning, nbi.J, qj);
                E3 f[3];
                double d[3];
                faceFlux(pr.pvt, qi, qj, I, nbi.J, nbi, pr.G.volume(I), pr.G.volume(nbi.J), f, d, pr.diffusion(),
                         pr.dispersion(), dsp ? nv[I] : 0.0, dsp ? nv[nbi.J] : 0.0, pr.rock.tab, pr.P->dir_mob,
                         pr.G.N);
                for (int e = 0; e < 3; ++e) {
                    if (flows) flows[3L * k + e] = f[e].v;
                    if (flores) flores[3L * k + e] = d[e];
                }
            }
        }
        return 0;
    } catch (...) { return -1; }
}

int ob_linearize_domain(const ob_problem* p, const double* pv, const uint8_t* meaning, const double* storage1,
                        double dt, int32_t n, const int32_t* cells, double* residual, double* jac)
{
    try {
        Problem pr(p);
        if (n <= 0 || !cells) return -1;
        std::vector<uint8_t> dom(pr.G.N, 0);
        for (int32_t k = 0; k < n; ++k) {
            if (cells[k] < 0 || cells[k] >= pr.G.N || dom[cells[k]]) return -1;
            dom[cells[k]] = 1;
        }
        const long R = (long)pr.pat.rowPtr.size() - 1;
        linearize(pr, pv, meaning, storage1, dt, residual, jac, 1, nullptr, nullptr, n == R ? nullptr : dom.data());
        return 0;
    } catch (...) { return -1; }
}

int ob_linearize(const ob_problem* p, const double* pv, const uint8_t* meaning,
                 const double* storage1, double dt, double* residual, double* jac, int nthreads)
{
    return ob_linearize_ext(p, nullptr, nullptr, nullptr, nullptr, pv, meaning, storage1, dt, residual, jac,
                            nthreads);
}

int ob_linearize_ext(const ob_problem* p, const ob_diffusion* dif, const ob_dispersion* dsp, const double* nv_in,
                     double* nv_out, const double* pv, const uint8_t* meaning, const double* storage1, double dt,
                     double* residual, double* jac, int nthreads)
{
    try {
        Problem pr(p, dif, dsp);
        if (dsp && (!nv_in || !nv_out)) throw std::invalid_argument("dispersion needs the cells' velocities");
        linearize(pr, pv, meaning, storage1, dt, residual, jac, nthreads, nv_in, nv_out);
        return 0;
    } catch (...) { return -1; }
}

/* The CPU baseline's assembly figure: the problem and its pattern built once,
 * then reps linearizations on nthreads threads; seconds[k] the wall time of
 * the k-th (the loop of linearize alone). */
int ob_linearize_timed(const ob_problem* p, const double* pv, const uint8_t* meaning, const double* storage1,
                       double dt, int nthreads, int reps, double* seconds)
{
    try {
        Problem pr(p);
        const long R = (long)pr.pat.rowPtr.size() - 1, nnzb = pr.pat.rowPtr[R];
        std::vector<double> r(3 * R), J(9 * nnzb);
        for (int k = 0; k < reps; ++k) {
            const auto t0 = std::chrono::steady_clock::now();
            linearize(pr, pv, meaning, storage1, dt, r.data(), J.data(), nthreads);
            seconds[k] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        return 0;
    } catch (...) { return -1; }
}

int ob_ilu0_factor(int n, int bs, const int32_t* row_ptr, const int32_t* col_idx, double* vals)
{
    return ilu0Factor<double>(n, bs, row_ptr, col_idx, vals);
}

int ob_ilu0_apply(int n, int bs, const int32_t* row_ptr, const int32_t* col_idx,
                  const double* lu, const double* d, double* v, double relax)
{
    ilu0Apply<double>(n, bs, row_ptr, col_idx, lu, d, v, relax);
    return 0;
}

int ob_jacobi_apply(int n, int bs, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                    const double* d, double* v, int order, double relax)
{
    /* the preconditioner of ob_bicgstab with precond = 1 applied to v in place */
    const int bb = bs * bs;
    std::vector<double> M((size_t)bb * n);
    for (int i = 0; i < n; ++i) {
        int k = row_ptr[i];
        while (col_idx[k] != i) ++k;
        std::copy(vals + (long)bb * k, vals + (long)bb * (k + 1), M.begin() + (long)bb * i);
        if (!invertBlock(bs, M.data() + (long)bb * i)) return 2;
    }
    Precond pc{1, bs, n, relax, order, row_ptr, col_idx};
    std::vector<double> in(d, d + (long)n * bs), out(v, v + (long)n * bs);
    jacobiSweeps<double>(pc, vals, M.data(), in, out);
    std::copy(out.begin(), out.end(), v);
    return 0;
}

int ob_colored_ilu(int nx, int ny, int nz, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                   const double* d, double* v, double relax)
{
    const TileColoring tc(nx, ny, nz);
    const long n = (long)nx * ny * nz;
    std::vector<double> dinv(9 * n);
    if (coloredIluFactor(tc, row_ptr, col_idx, vals, dinv.data()) != 0) return 2;
    coloredIluApply(tc, row_ptr, col_idx, vals, dinv.data(), d, v, relax);
    return 0;
}

int ob_bicgstab_grid(int nx, int ny, int nz, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                     const double* b, double* x, const ob_params* prm, int32_t* iterations)
{
    const int dims[3] = {nx, ny, nz};
    int it = 0;
    const int rc = backendSolve<double>(nx * ny * nz, 3, row_ptr, col_idx, vals, b, x, prm, &it, dims);
    *iterations = it;
    return rc;
}

/* the AMG hierarchy's level dimensions ([levels][3]); returns the number of
 * levels (at most cap), -1 for invalid parameters */
int ob_amg_dims(int nx, int ny, int nz, const ob_params* prm, int32_t* dims, int cap)
{
    if (!amgParamsValid(prm)) return -1;
    const auto d = amgDims(nx, ny, nz, prm->amg_agg, prm->amg_coarsen_target);
    for (size_t l = 0; l < d.size() && (int)l < cap; ++l)
        for (int q = 0; q < 3; ++q) dims[3 * l + q] = d[l][q];
    return (int)d.size();
}

/* level `level`'s matrix in the structured layout [7][N_l][9] */
int ob_amg_level_matrix(int nx, int ny, int nz, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                        const ob_params* prm, int level, double* A7)
{
    if (!amgParamsValid(prm)) return -1;
    Amg amg;
    const int rc = amgSetup(amg, nx, ny, nz, diaFromBcrs(nx, ny, nz, row_ptr, col_idx, vals), prm->amg_agg,
                            prm->amg_coarsen_target);
    if (level < 0 || level >= (int)amg.lv.size()) return -1;
    std::copy(amg.lv[level].A.begin(), amg.lv[level].A.end(), A7);
    return rc;
}

/* v = one V-cycle applied to d */
int ob_amg_apply(int nx, int ny, int nz, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                 const ob_params* prm, const double* d, double* v)
{
    if (!amgParamsValid(prm)) return -1;
    Amg amg;
    amg.pre = prm->amg_pre_smooth; amg.post = prm->amg_post_smooth; amg.damp = prm->amg_prolong_damp;
    if (amgSetup(amg, nx, ny, nz, diaFromBcrs(nx, ny, nz, row_ptr, col_idx, vals), prm->amg_agg,
                 prm->amg_coarsen_target) != 0)
        return 2;
    amgCycle(amg, 0, d, v);
    return 0;
}

/* one red-black Gauss-Seidel sweep of the structured-layout matrix A7 in place */
int ob_gs_sweep(int nx, int ny, int nz, const double* A7, const double* b, double* x, int forward, int zero)
{
    AmgLevel L;
    L.nx = nx; L.ny = ny; L.nz = nz; L.N = (long)nx * ny * nz;
    L.A.assign(A7, A7 + 63 * L.N);
    L.dinv.resize(9 * L.N);
    for (long I = 0; I < L.N; ++I) {
        std::copy(A7 + 9 * (3 * L.N + I), A7 + 9 * (3 * L.N + I) + 9, L.dinv.begin() + 9 * I);
        if (!invertBlock(3, L.dinv.data() + 9 * I)) return 2;
    }
    amgGsSweep(L, b, x, forward != 0, zero != 0);
    return 0;
}

/* in-place inverse of a dense n x n matrix (row-major) */
int ob_dense_inverse(int n, double* M)
{
    return denseInverse(n, M);
}

double ob_dot(int64_t m, const double* x, const double* y)
{
    std::vector<double> a(x, x + m), b(y, y + m);
    return dot<double>(a, b);
}

int ob_spmv(int n, int bs, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
            const double* x, double* y)
{
    spmv<double>(n, bs, row_ptr, col_idx, vals, x, y);
    return 0;
}

int ob_bicgstab(int n, int bs, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                const double* b, double* x, const ob_params* prm, int32_t* iterations)
{
    int it = 0, rc;
    if (prm->solver_float) rc = backendSolve<float>(n, bs, row_ptr, col_idx, vals, b, x, prm, &it);
    else rc = backendSolve<double>(n, bs, row_ptr, col_idx, vals, b, x, prm, &it);
    *iterations = it;
    return rc;
}

int ob_update_primary(const ob_problem* p, const double* cur, const uint8_t* cur_meaning,
                      const double* delta, double* next, uint8_t* next_meaning,
                      uint8_t* was_switched, const ob_params* prm, int32_t* num_switched)
{
    Pvt pvt; pvt.init(p);
    const long N = (long)p->nx * p->ny * p->nz;
    for (long k = 0; k < 3 * N; ++k)
        if (!std::isfinite(delta[k])) return 2;
    int sw = 0;
    for (long I = 0; I < N; ++I) {
        updatePrimary(pvt, cur + 3 * I, cur_meaning[I], delta + 3 * I, next + 3 * I,
                      next_meaning[I], was_switched[I], prm);
        sw += was_switched[I];
    }
    *num_switched = sw;
    return 0;
}

/* Simulator::run (simulator.hh:697-810) -> FvBaseProblem::timeIntegration
 * (fvbaseproblem.hh:512-557) -> NewtonMethod::apply (newtonmethod.hh:228-438),
 * dt control by suggestTimeStepSize (newtonmethod.hh:448-465). */
int ob_simulate(const ob_problem* p, double* pv, uint8_t* meaning, int n_steps, double dt0,
                const ob_params* prm, double* pv_hist, double* dt_hist, ob_iter_log* log,
                int log_cap, int32_t* log_len, double* timers, int nthreads)
{
    return ob_simulate_ext(p, nullptr, nullptr, pv, meaning, n_steps, dt0, prm, pv_hist, dt_hist, log, log_cap,
                           log_len, timers, nthreads);
}

/* NewtonMethod::apply inside FvBaseProblem::timeIntegration; with auxiliary
 * DOFs (aux: their solution(0), [n_aux][3]) the module's terms come from fn
 * after every linearization (linearizeAuxiliaryEquations_, newtonmethod.hh:
 * 287) */
static int simulateImpl(const ob_problem* p, const ob_diffusion* dif, const ob_dispersion* dsp, double* pv,
                 uint8_t* meaning, double* aux, ob_aux_fn fn, void* user, int n_steps, double dt0,
                 const ob_params* prm, double* pv_hist, double* dt_hist, ob_iter_log* log, int log_cap,
                 int32_t* log_len, double* timers, int nthreads)
{
    using clk = std::chrono::steady_clock;
    try {
        if (p->n_aux > 0 && (!fn || !aux)) throw std::invalid_argument("auxiliary equations need their module");
        Problem pr(p, dif, dsp);
        /* the cells' oil velocities of the last linearization, zeros before the first */
        std::vector<double> nvPrev(dsp ? pr.G.N : 0, 0.0), nvNext(dsp ? pr.G.N : 0, 0.0);
        const long N = pr.G.N;
        const long nAux = p->n_aux > 0 ? p->n_aux : 0, R = N + nAux;
        const int32_t* rp = pr.pat.rowPtr.data();
        const int32_t* ci = pr.pat.colIdx.data();
        const long nnzb = rp[R];
        std::vector<double> sol1(pv, pv + 3 * N), st1(3 * N), res(3 * R), jac(9 * nnzb), upd(3 * R),
            cur(3 * N), aux1(aux ? aux : pv, aux ? aux + 3 * nAux : pv);
        std::vector<uint8_t> mean1(meaning, meaning + N), curMean(N), wasSw(N, 0);
        double tLin = 0, tSol = 0, tUpd = 0;
        auto t0 = clk::now();
        int nlog = 0;
        double dt = dt0;
        for (int step = 0; step < n_steps; ++step) {
            bool stepOk = false;
            int numIter = 0;
            for (int attempt = 0; attempt < prm->max_dt_divisions; ++attempt) {
                /* NewtonMethod::apply */
                numIter = 0;
                double error = 1e100, lastError = 1e100;
                bool failed = false;
                auto converged = [&]() { return error <= prm->newton_tol; };
                auto proceed = [&]() {
                    if (numIter < 1) return true;
                    if (converged()) return false;
                    if (numIter >= prm->newton_max_iter) return error * 4.0 < lastError;
                    return true;
                };
                while (proceed()) {
                    std::copy(pv, pv + 3 * N, cur.begin());
                    std::copy(meaning, meaning + N, curMean.begin());
                    auto a = clk::now();
                    if (numIter == 0) computeStorageAll(pr, cur.data(), curMean.data(), st1.data());
                    linearize(pr, cur.data(), curMean.data(), st1.data(), dt, res.data(), jac.data(), nthreads,
                              nvPrev.data(), nvNext.data());
                    nvPrev.swap(nvNext);
                    if (nAux > 0 && fn(user, cur.data(), curMean.data(), aux, jac.data(), res.data(), numIter) != 0) {
                        failed = true;
                        break;
                    }
                    auto b = clk::now();
                    tLin += std::chrono::duration<double>(b - a).count();
                    /* preSolve_ */
                    lastError = error;
                    error = 0;
                    for (long k = 0; k < 3 * N; ++k) error = std::max(std::abs(res[k]), error);
                    if (error > prm->newton_max_error) { failed = true; break; }
                    if (!proceed()) {
                        if (nlog < log_cap && log) log[nlog++] = {step, attempt, numIter, 0, 0, 1, error, dt};
                        ++numIter;
                        break;
                    }
                    auto c = clk::now();
                    int it = 0;
                    int rc;
                    const int dims[3] = {p->nx, p->ny, p->nz};
                    if (prm->solver_float) rc = backendSolve<float>((int)R, 3, rp, ci, jac.data(), res.data(), upd.data(), prm, &it, dims);
                    else rc = backendSolve<double>((int)R, 3, rp, ci, jac.data(), res.data(), upd.data(), prm, &it, dims);
                    auto d = clk::now();
                    tSol += std::chrono::duration<double>(d - c).count();
                    if (rc != 0) {
                        if (nlog < log_cap && log) log[nlog++] = {step, attempt, numIter, it, 0, 0, error, dt};
                        failed = true;
                        break;
                    }
                    /* update_ */
                    bool finite = true;
                    for (long k = 0; k < 3 * R; ++k) if (!std::isfinite(upd[k])) { finite = false; break; }
                    if (!finite) { failed = true; break; }
                    int nsw = 0;
                    for (long I = 0; I < N; ++I) {
                        updatePrimary(pr.pvt, cur.data() + 3 * I, curMean[I], upd.data() + 3 * I,
                                      pv + 3 * I, meaning[I], wasSw[I], prm);
                        nsw += wasSw[I];
                    }
                    /* the auxiliary DOFs: next = current - update (newtonmethod.hh:704-708) */
                    for (long k = 0; k < 3 * nAux; ++k) aux[k] = aux[k] - upd[3 * N + k];
                    tUpd += std::chrono::duration<double>(clk::now() - d).count();
                    if (nlog < log_cap && log) log[nlog++] = {step, attempt, numIter, it, nsw, 1, error, dt};
                    if (prm->verbose)
                        printf("step %d attempt %d iter %d error %.6e lin %d switched %d\n", step, attempt,
                               numIter, error, it, nsw);
                    ++numIter;
                }
                if (!failed && converged()) { stepOk = true; break; }
                /* failed_: numIterations = 2 * target; updateFailed: restore */
                numIter = 2 * prm->newton_target_iter;
                std::copy(sol1.begin(), sol1.end(), pv);
                std::copy(mean1.begin(), mean1.end(), meaning);
                if (nAux > 0) std::copy(aux1.begin(), aux1.end(), aux);
                double ndt = dt / 2.0;
                if (dt < prm->min_dt * (1 + 1e-9)) break;
                if (ndt < prm->min_dt) ndt = prm->min_dt;
                dt = ndt;
            }
            if (!stepOk) { if (log_len) *log_len = nlog; return 3; }
            if (dt_hist) dt_hist[step] = dt;
            /* advanceTimeLevel */
            std::copy(pv, pv + 3 * N, sol1.begin());
            std::copy(meaning, meaning + N, mean1.begin());
            if (nAux > 0) std::copy(aux, aux + 3 * nAux, aux1.begin());
            if (pv_hist) std::copy(pv, pv + 3 * N, pv_hist + (long)step * 3 * N);
            /* suggestTimeStepSize */
            const int tgt = prm->newton_target_iter;
            double nd;
            if (numIter > tgt) {
                double percent = double(numIter - tgt) / tgt;
                nd = std::max(prm->min_dt, dt / (1.0 + percent));
            } else {
                double percent = double(tgt - numIter) / tgt;
                nd = std::max(prm->min_dt, dt * (1.0 + percent / 1.2));
            }
            dt = std::min(prm->max_dt, nd);
        }
        if (log_len) *log_len = nlog;
        if (timers) {
            timers[0] = tLin; timers[1] = tSol; timers[2] = tUpd;
            timers[3] = std::chrono::duration<double>(clk::now() - t0).count();
        }
        return 0;
    } catch (const std::exception& e) {
        fprintf(stderr, "ob_simulate: %s\n", e.what());
        return -1;
    }
}

int ob_simulate_ext(const ob_problem* p, const ob_diffusion* dif, const ob_dispersion* dsp, double* pv,
                    uint8_t* meaning, int n_steps, double dt0, const ob_params* prm, double* pv_hist,
                    double* dt_hist, ob_iter_log* log, int log_cap, int32_t* log_len, double* timers, int nthreads)
{
    return simulateImpl(p, dif, dsp, pv, meaning, nullptr, nullptr, nullptr, n_steps, dt0, prm, pv_hist, dt_hist,
                        log, log_cap, log_len, timers, nthreads);
}

int ob_simulate_aux(const ob_problem* p, double* pv, uint8_t* meaning, double* aux, ob_aux_fn fn, void* user,
                    int n_steps, double dt0, const ob_params* prm, double* pv_hist, double* dt_hist,
                    ob_iter_log* log, int log_cap, int32_t* log_len, int nthreads)
{
    return simulateImpl(p, nullptr, nullptr, pv, meaning, aux, fn, user, n_steps, dt0, prm, pv_hist, dt_hist, log,
                        log_cap, log_len, nullptr, nthreads);
}

} // extern "C"
