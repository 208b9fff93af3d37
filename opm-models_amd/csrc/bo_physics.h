// bo_physics.h -- device-side black-oil physics of the TPFA hot path.
//
// Restates, for the water-meaning Sw / pressure-meaning Po / gas-meaning
// {Sg, Rs} subset used by the black-oil configurations:
//   BlackOilIntensiveQuantities::update      blackoilintensivequantities.hh:179-516
//   BlackOilLocalResidualTPFA::computeStorage blackoillocalresidualtpfa.hh:151-227,824-851
//   BlackOilLocalResidualTPFA::calculateFluxes_ :325-504 with evalPhaseFluxes_ :752-811
//   pressure difference and upwinding        transfluxmodule.hh:254-303
//
// Operation order is fixed (the library is built with -ffp-contract=off) so
// that the FP64 residual and Jacobian are reproducible to the last bit by a
// sequential CPU evaluation of the same formulas.
//
// Forms chosen where the arithmetic lives outside opm-models (opm-common's
// DenseAd, LinearMaterial, LiveOil/DryGas/ConstantCompressibilityWater PVT):
//   a/b      = a * (1/b);  (a/b)' = (a' - (a/b) b') * (1/b)
//   min/max  keep the differentiated operand at ties
//   kr       = max(min(S, 1), 0), pc = 0
//   1/B, mu  piecewise linear in pressure, linear extrapolation at both ends
//   undersaturated oil: pb = Rs_sat^-1(Rs), 1/Bo = 1/Bo_sat(pb) (1 + c_o (p - pb)),
//                       mu_o = mu_o_sat(pb)
//   water:   1/Bw = 1 + X (1 + X/2), X = c_w (p - p_ref), constant mu_w
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace omodels {

constexpr int kMaxTab = 12;

// One piecewise-linear table, passed by value in the kernel argument block:
// all lanes read the same nodes, the segment is selected without indexing.
struct Tab {
    int n;
    double x[kMaxTab], y[kMaxTab], s[kMaxTab];
};

struct BoFluid {
    Tab rsSat, pOfRs, invBoSat, muoSat, invBg, mug;
    double co, pref, cw, muw;
    double rho[3];          // surface densities water, oil, gas
};

enum : int { kW = 0, kO = 1, kG = 2 };          // phases
enum : int { kCO = 0, kCW = 1, kCG = 2 };       // components / equations
enum : int { kMeanSg = 0, kMeanRs = 1 };        // gas meaning

__host__ __device__ inline int compOfPhase(int a) { return a == kW ? kCW : (a == kO ? kCO : kCG); }

// ---------------------------------------------------------------------------
// Forward-mode AD with three directional derivatives (Sw, po, Sg|Rs).
struct E3 {
    double v, d0, d1, d2;
};

__device__ __forceinline__ E3 ec(double x) { return E3{x, 0.0, 0.0, 0.0}; }
__device__ __forceinline__ E3 evar(double x, int idx)
{
    return E3{x, idx == 0 ? 1.0 : 0.0, idx == 1 ? 1.0 : 0.0, idx == 2 ? 1.0 : 0.0};
}
__device__ __forceinline__ E3 operator+(E3 a, E3 b) { return E3{a.v + b.v, a.d0 + b.d0, a.d1 + b.d1, a.d2 + b.d2}; }
__device__ __forceinline__ E3 operator-(E3 a, E3 b) { return E3{a.v - b.v, a.d0 - b.d0, a.d1 - b.d1, a.d2 - b.d2}; }
__device__ __forceinline__ E3 operator*(E3 a, E3 b)
{
    return E3{a.v * b.v, a.d0 * b.v + a.v * b.d0, a.d1 * b.v + a.v * b.d1, a.d2 * b.v + a.v * b.d2};
}
__device__ __forceinline__ E3 operator/(E3 a, E3 b)
{
    const double inv = 1.0 / b.v;
    const double v = a.v * inv;
    return E3{v, (a.d0 - v * b.d0) * inv, (a.d1 - v * b.d1) * inv, (a.d2 - v * b.d2) * inv};
}
__device__ __forceinline__ E3 operator+(E3 a, double s) { return E3{a.v + s, a.d0, a.d1, a.d2}; }
__device__ __forceinline__ E3 operator-(E3 a, double s) { return E3{a.v - s, a.d0, a.d1, a.d2}; }
__device__ __forceinline__ E3 operator-(double s, E3 a) { return E3{s - a.v, -a.d0, -a.d1, -a.d2}; }
__device__ __forceinline__ E3 operator*(E3 a, double s) { return E3{a.v * s, a.d0 * s, a.d1 * s, a.d2 * s}; }
__device__ __forceinline__ E3 operator-(E3 a) { return E3{-a.v, -a.d0, -a.d1, -a.d2}; }
__device__ __forceinline__ E3 emin(E3 a, double c) { return a.v <= c ? a : ec(c); }
__device__ __forceinline__ E3 emax(E3 a, double c) { return a.v >= c ? a : ec(c); }

// Segment selection: the largest i in [0, n-2] with x_i <= x (0 below x_1).
__device__ __forceinline__ void tabSeg(const Tab& t, double x, double& x0, double& y0, double& s0)
{
    x0 = t.x[0]; y0 = t.y[0]; s0 = t.s[0];
#pragma unroll
    for (int i = 1; i < kMaxTab - 1; ++i) {
        if (i + 1 < t.n && x >= t.x[i]) { x0 = t.x[i]; y0 = t.y[i]; s0 = t.s[i]; }
    }
}
__device__ __forceinline__ double tabEval(const Tab& t, double x)
{
    double x0, y0, s0;
    tabSeg(t, x, x0, y0, s0);
    return y0 + s0 * (x - x0);
}
__device__ __forceinline__ E3 tabEval(const Tab& t, E3 x)
{
    double x0, y0, s0;
    tabSeg(t, x.v, x0, y0, s0);
    return E3{y0 + s0 * (x.v - x0), s0 * x.d0, s0 * x.d1, s0 * x.d2};
}

// ---------------------------------------------------------------------------
// Intensive quantities.  The value-only variant feeds the neighbour side of
// a face; it performs the value half of every AD operation above, so both
// sides of a face see identical bits.
struct IqV {                 // neighbour (values)
    double p, lam[3], invB[3], rho[3], Rs;
};
struct IqE {                 // own cell (with derivatives)
    E3 S[3], p, lam[3], invB[3], rho[3], Rs;
};

__device__ __forceinline__ void iqValues(const BoFluid& f, double sw, double po, double x2, int mean, IqV& q)
{
    const double sg = (mean == kMeanSg) ? x2 : 0.0;
    const double so = (1.0 - sw) - sg;
    const double S[3] = {sw, so, sg};
    q.p = po;
    q.Rs = (mean == kMeanRs) ? x2 : tabEval(f.rsSat, po);
    double mu[3];
    {
        const double X = (po - f.pref) * f.cw;
        const double u = X * 0.5 + 1.0;
        q.invB[kW] = X * u + 1.0;
        mu[kW] = f.muw;
    }
    if (mean == kMeanSg) {
        q.invB[kO] = tabEval(f.invBoSat, po);
        mu[kO] = tabEval(f.muoSat, po);
    } else {
        const double pb = tabEval(f.pOfRs, q.Rs);
        const double fac = (po - pb) * f.co + 1.0;
        q.invB[kO] = tabEval(f.invBoSat, pb) * fac;
        mu[kO] = tabEval(f.muoSat, pb);
    }
    q.invB[kG] = tabEval(f.invBg, po);
    mu[kG] = tabEval(f.mug, po);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double kr = S[a] <= 1.0 ? S[a] : 1.0;
        kr = kr >= 0.0 ? kr : 0.0;
        q.lam[a] = kr * (1.0 / mu[a]);
    }
    q.rho[kW] = q.invB[kW] * f.rho[kW];
    q.rho[kG] = q.invB[kG] * f.rho[kG];
    q.rho[kO] = q.invB[kO] * f.rho[kO] + (q.invB[kO] * q.Rs) * f.rho[kG];
}

__device__ __forceinline__ void iqEval(const BoFluid& f, double sw, double po, double x2, int mean, IqE& q)
{
    const E3 Sw = evar(sw, 0);
    const E3 Sg = (mean == kMeanSg) ? evar(x2, 2) : ec(0.0);
    const E3 So = (1.0 - Sw) - Sg;
    q.S[kW] = Sw; q.S[kO] = So; q.S[kG] = Sg;
    const E3 p = evar(po, 1);
    q.p = p;
    q.Rs = (mean == kMeanRs) ? evar(x2, 2) : tabEval(f.rsSat, p);
    E3 mu[3];
    {
        const E3 X = (p - f.pref) * f.cw;
        const E3 u = X * 0.5 + 1.0;
        q.invB[kW] = X * u + 1.0;
        mu[kW] = ec(f.muw);
    }
    if (mean == kMeanSg) {
        q.invB[kO] = tabEval(f.invBoSat, p);
        mu[kO] = tabEval(f.muoSat, p);
    } else {
        const E3 pb = tabEval(f.pOfRs, q.Rs);
        const E3 fac = (p - pb) * f.co + 1.0;
        q.invB[kO] = tabEval(f.invBoSat, pb) * fac;
        mu[kO] = tabEval(f.muoSat, pb);
    }
    q.invB[kG] = tabEval(f.invBg, p);
    mu[kG] = tabEval(f.mug, p);
#pragma unroll
    for (int a = 0; a < 3; ++a) q.lam[a] = emax(emin(q.S[a], 1.0), 0.0) / mu[a];
    q.rho[kW] = q.invB[kW] * f.rho[kW];
    q.rho[kG] = q.invB[kG] * f.rho[kG];
    q.rho[kO] = q.invB[kO] * f.rho[kO] + (q.invB[kO] * q.Rs) * f.rho[kG];
}

// computeStorage: surface volumes times surface densities.
__device__ __forceinline__ void storageEval(const BoFluid& f, const IqE& q, double phi, E3 st[3])
{
    st[0] = ec(0.0); st[1] = ec(0.0); st[2] = ec(0.0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const E3 sv = (q.S[a] * q.invB[a]) * phi;
        const int c = compOfPhase(a);
        st[c] = st[c] + sv;
        if (a == kO) st[kCG] = st[kCG] + q.Rs * sv;
    }
    st[kCW] = st[kCW] * f.rho[kW];
    st[kCG] = st[kCG] * f.rho[kG];
    st[kCO] = st[kCO] * f.rho[kO];
}

// the same with a porosity that depends on the cell's pressure (rock compressibility)
__device__ __forceinline__ void storageEval(const BoFluid& f, const IqE& q, E3 phi, E3 st[3])
{
    st[0] = ec(0.0); st[1] = ec(0.0); st[2] = ec(0.0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const E3 sv = (q.S[a] * q.invB[a]) * phi;
        const int c = compOfPhase(a);
        st[c] = st[c] + sv;
        if (a == kO) st[kCG] = st[kCG] + q.Rs * sv;
    }
    st[kCW] = st[kCW] * f.rho[kW];
    st[kCG] = st[kCG] * f.rho[kG];
    st[kCO] = st[kCO] * f.rho[kO];
}

__device__ __forceinline__ void storageValues(const BoFluid& f, double sw, double po, double x2, int mean,
                                              double phi, double st[3])
{
    IqV q;
    iqValues(f, sw, po, x2, mean, q);
    const double sg = (mean == kMeanSg) ? x2 : 0.0;
    const double so = (1.0 - sw) - sg;
    const double S[3] = {sw, so, sg};
    st[0] = 0.0; st[1] = 0.0; st[2] = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double sv = (S[a] * q.invB[a]) * phi;
        const int c = compOfPhase(a);
        st[c] = st[c] + sv;
        if (a == kO) st[kCG] = st[kCG] + q.Rs * sv;
    }
    st[kCW] = st[kCW] * f.rho[kW];
    st[kCG] = st[kCG] * f.rho[kG];
    st[kCO] = st[kCO] * f.rho[kO];
}

// Molecular diffusion in the oil phase (BlackOilDiffusionModule).  toFrac is
// toFractionGasOil (blackoildiffusionmodule.hh:242-248), rho_o M_g / (rho_g M_o)
// with unit molar masses unless the mole-fraction form is asked for; effOil
// and effGas are the effective coefficients of the oil and the gas component
// in the oil phase: the arithmetic mean of the two cells' constant
// coefficients (:568-590), tortuosity being disabled for black-oil (:384-394).
// diff: the problem's diffusivity per face, [3][N] like the transmissibilities.
// the two cells' transmissibility multipliers of one face
struct FaceTM {
    E3 in;
    double ex;
};

// Directional relative permeability multipliers of a face's two cells for
// the face's axis, per phase: the upwind cell's mobility along the face is
// its mobility times its multiplier (BlackOilIntensiveQuantities::mobility
// (phaseIdx, facedir), blackoilintensivequantities.hh:530-550, with the
// problem's directional relative permeabilities from updateRelperms,
// multiphasebaseproblem.hh:257-261, taken as kr_dir = m_dir kr).
struct FaceDM {
    double in[3], ex[3];
};

// Rock compressibility and compaction (blackoilintensivequantities.hh:447-480):
// porosity *= 1 + x + x^2/2 with x = c (p_o - p_ref) where c > 0, then times
// the problem's pore-volume multiplier of p_o (rockCompPoroMultiplier); the
// face term takes the transmissibility multiplier of p_o
// (rockCompTransMultiplier).  comp / pref per cell (nullptr: 0 / 1e5 Pa);
// poroMult / transMult: piecewise-linear tables of p_o, or nullptr (1).
struct BoRock {
    const double* comp;
    const double* pref;
    const Tab* poroMult;
    const Tab* transMult;
};

__device__ __forceinline__ E3 rockPorosity(const BoRock& r, long I, double poro, E3 p)
{
    E3 phi = ec(poro);
    const double c = r.comp ? r.comp[I] : 0.0;
    if (c > 0.0) {
        const double pref = r.pref ? r.pref[I] : 1e5;
        const E3 x = (p - pref) * c;
        phi = phi * ((x + 1.0) + (x * 0.5) * x);
    }
    if (r.poroMult) phi = phi * tabEval(*r.poroMult, p);
    return phi;
}
__device__ __forceinline__ double rockPorosityV(const BoRock& r, long I, double poro, double p)
{
    double phi = poro;
    const double c = r.comp ? r.comp[I] : 0.0;
    if (c > 0.0) {
        const double pref = r.pref ? r.pref[I] : 1e5;
        const double x = (p - pref) * c;
        phi = phi * ((x + 1.0) + (x * 0.5) * x);
    }
    if (r.poroMult) phi = phi * tabEval(*r.poroMult, p);
    return phi;
}
__device__ __forceinline__ E3 rockTransMult(const BoRock& r, E3 p) { return r.transMult ? tabEval(*r.transMult, p) : ec(1.0); }
__device__ __forceinline__ double rockTransMultV(const BoRock& r, double p) { return r.transMult ? tabEval(*r.transMult, p) : 1.0; }

struct BoDiff {
    double toFrac, effOil, effGas;
    const double* diff;
};
// Dispersion in the oil phase (BlackOilDispersionModule): toMassFrac is
// rho_o / rho_g (toMassFractionGasOil, blackoildispersionmodule.hh:258-262);
// disp the problem's dispersivity per face ([3][N] like the
// transmissibilities); nvIn each cell's oil normVelocityCell of the previous
// linearization, nvOut where this linearization writes its own.
struct BoDisp {
    double toMassFrac;
    const double* disp;
    double* nvIn;            // read only by the kernels
    double* nvOut;
};
// the neighbour's side of one face: So * 1/Bo and the face's diffusivity;
// the face's dispersivity and the mean of the two cells' oil velocities
struct DiffFace {
    double sbEx, diffusivity, dispersivity, nvAvg;
};

// One face seen from its interior cell I (derivatives w.r.t. I only), times
// the face area.  upInTie: upwind choice when the pressure difference is
// exactly zero (equal volumes -> the smaller global index, transfluxmodule.hh:283-301).
// WithDarcy: also the phases' reservoir-volume rates across the face,
// darcyFlux.value() * faceArea at the phase's component index (the FLORES
// values, blackoillocalresidualtpfa.hh:376-385)
// WithDiff / WithDisp: plus the diffusive / dispersive term of the oil phase.
// tm: nullptr, or the two cells' rockCompTransMultiplier (the interior's with
// derivatives, the exterior's value): the face takes their arithmetic mean
// (blackoillocalresidualtpfa.hh:372-381).
template <bool WithDarcy, bool WithDiff = false, bool WithDisp = false>
__device__ __forceinline__ void faceFluxT(const BoFluid& f, const IqE& in, const IqV& ex, double trans,
                                          double area, double dZg, bool upInTie, E3 flux[3], double darcyOut[3],
                                          const BoDiff* dif = nullptr, DiffFace df = DiffFace{0.0, 0.0, 0.0, 0.0},
                                          const BoDisp* dsp = nullptr, double thp = 0.0, const FaceTM* tm = nullptr,
                                          const FaceDM* dm = nullptr)
{
    flux[0] = ec(0.0); flux[1] = ec(0.0); flux[2] = ec(0.0);
    if constexpr (WithDarcy) { darcyOut[0] = 0.0; darcyOut[1] = 0.0; darcyOut[2] = 0.0; }
    const double tf = (-trans) / area;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (in.lam[a].v <= 0.0 && ex.lam[a] <= 0.0) continue;
        const E3 rhoAvg = (in.rho[a] + ex.rho[a]) * 0.5;
        const E3 pEx = rhoAvg * dZg + ex.p;
        E3 dp = pEx - in.p;
        // threshold pressure of the face, in the form of the reference's
        // solvent flux (blackoilsolventmodules.hh:1484-1492): below it the
        // face carries no flow
        if (thp > 0.0) {
            if (fabs(dp.v) > thp) dp = dp.v < 0.0 ? dp + thp : dp - thp;
            else dp = ec(0.0);
        }
        const bool upIn = dp.v > 0.0 ? false : (dp.v < 0.0 ? true : upInTie);
        E3 svf;
        E3 gas;
        double darcyV;
        if (upIn) {
            const E3 lam = dm ? in.lam[a] * dm->in[a] : in.lam[a];
            const E3 darcy = tm ? ((dp * lam) * ((tm->in + tm->ex) * 0.5)) * tf : (dp * lam) * tf;
            svf = in.invB[a] * darcy;
            if (a == kO) gas = in.Rs * svf;
            darcyV = darcy.v;
        } else {
            const double lam = dm ? ex.lam[a] * dm->ex[a] : ex.lam[a];
            const E3 darcy = tm ? dp * ((((tm->in + tm->ex) * 0.5) * lam) * tf) : dp * (lam * tf);
            svf = darcy * ex.invB[a];
            if (a == kO) gas = svf * ex.Rs;
            darcyV = darcy.v;
        }
        const int c = compOfPhase(a);
        flux[c] = flux[c] + svf * f.rho[a];
        if (a == kO) flux[kCG] = flux[kCG] + gas * f.rho[kG];
        if constexpr (WithDarcy) darcyOut[c] = darcyV * area;
    }
    // BlackOilDiffusionModule::addDiffusiveFlux (blackoildiffusionmodule.hh:168-238)
    // after the phase loop, with the face's diffusivity per area
    // (blackoillocalresidualtpfa.hh:469-479).  Only the oil phase has a term:
    // water is skipped (no gas in water, :180-183) and the gas phase, without
    // vaporized oil, has diffR = 0, so its terms are exact zeros (:205-209).
    // The terms enter in surface-volume units, as the reference adds them.
    if constexpr (WithDiff) {
        E3 bSAvg = in.S[kO] * in.invB[kO];
        bSAvg = bSAvg + df.sbEx;
        bSAvg = bSAvg * 0.5;
        if (!(bSAvg.v < 1.0e-6)) {
            const E3 rsAvg = (in.Rs + ex.Rs) * 0.5;
            const E3 convFactor = ec(1.0) / (rsAvg + dif->toFrac);
            const E3 diffR = in.Rs - ex.Rs;
            const double tmpDiff = df.diffusivity / area;
            // solvent (oil in oil) and solute (gas in oil)
            flux[kCO] = flux[kCO] + (((-bSAvg * convFactor) * diffR) * tmpDiff) * dif->effOil;
            flux[kCG] = flux[kCG] + (((bSAvg * diffR) * convFactor) * tmpDiff) * dif->effGas;
        }
    }
    // BlackOilDispersionModule::addDispersiveFlux (blackoildispersionmodule.hh:176-254)
    // after the diffusive term (blackoillocalresidualtpfa.hh:481-493): the oil
    // phase only (no gas in water; the gas phase is disabled, :193-197), with
    // the mean 1/Bo, the mean of the two cells' normVelocityCell (:484-506) and
    // the face's dispersivity per area.
    if constexpr (WithDisp) {
        E3 bAvg = in.invB[kO];
        bAvg = bAvg + ex.invB[kO];
        bAvg = bAvg * 0.5;
        const E3 rsAvg = (in.Rs + ex.Rs) * 0.5;
        const E3 convFactor = ec(1.0) / (rsAvg + dsp->toMassFrac);
        const E3 diffR = in.Rs - ex.Rs;
        const double tmpDisp = df.dispersivity / area;
        flux[kCO] = flux[kCO] + ((((-bAvg * df.nvAvg) * convFactor) * tmpDisp) * diffR);
        flux[kCG] = flux[kCG] + ((((bAvg * df.nvAvg) * convFactor) * tmpDisp) * diffR);
    }
    flux[0] = flux[0] * area; flux[1] = flux[1] * area; flux[2] = flux[2] * area;
}

// BlackOilLocalResidualTPFA::computeBoundaryFlux (blackoillocalresidualtpfa.hh:
// 507-620) of a free (or Dirichlet) boundary face, times the face area
// (tpfalinearizer.hh:849-865).  qb: the boundary's fluid state; per phase the
// trans flux module's boundary gradient (transfluxmodule.hh:323-409) per
// area: the boundary pressure plus the mean density times the depth
// difference to the face, the upstream side by that difference (ties: the
// cell), the upstream mobility; the flux leaves with the cell's 1/B and Rs
// when the boundary pressure is below the cell's, enters with the
// boundary's when it is above; evalPhaseFluxes_ (:753-811) with the
// reference densities, then adaptMassConservationQuantities_ (:825-851)
// multiplies by them once more, as the reference does.
__device__ __forceinline__ void boundaryFreeFlux(const BoFluid& f, const IqE& qi, const IqV& qb, double trans,
                                                 double area, double dZg, E3 out[3])
{
    const double tf = (-trans) / area;
    E3 bdy[3] = {ec(0.0), ec(0.0), ec(0.0)};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        E3 tmp[3] = {ec(0.0), ec(0.0), ec(0.0)};
        const E3 rhoAvg = (qi.rho[a] + qb.rho[a]) * 0.5;
        const E3 pEx = rhoAvg * dZg + qb.p;
        const E3 dp = pEx - qi.p;
        const E3 darcy = dp.v > 0.0 ? (dp * qb.lam[a]) * tf : (dp * qi.lam[a]) * tf;
        const int c = compOfPhase(a);
        if (qb.p < qi.p.v) {
            const E3 svf = qi.invB[a] * darcy;
            tmp[c] = tmp[c] + svf * f.rho[a];
            if (a == kO) tmp[kCG] = tmp[kCG] + (qi.Rs * svf) * f.rho[kG];
        } else if (qb.p > qi.p.v) {
            const E3 svf = darcy * qb.invB[a];
            tmp[c] = tmp[c] + svf * f.rho[a];
            if (a == kO) tmp[kCG] = tmp[kCG] + (svf * qb.Rs) * f.rho[kG];
        }
#pragma unroll
        for (int e = 0; e < 3; ++e) bdy[e] = bdy[e] + tmp[e];
    }
    bdy[kCW] = bdy[kCW] * f.rho[kW];
    bdy[kCG] = bdy[kCG] * f.rho[kG];
    bdy[kCO] = bdy[kCO] * f.rho[kO];
#pragma unroll
    for (int e = 0; e < 3; ++e) out[e] = bdy[e] * area;
}

// Well source terms per unit volume (ReservoirProblem-style problem.source()).
__device__ __forceinline__ void sourceEval(const BoFluid& f, const IqE& q, int type, double a, double b,
                                           double V, E3 src[3])
{
    src[0] = ec(0.0); src[1] = ec(0.0); src[2] = ec(0.0);
    if (type == 0) { src[kCW] = ec(a / V); return; }
    const E3 drive = q.p - b;
    if (drive.v <= 0.0) return;
    const double invV = 1.0 / V;
#pragma unroll
    for (int ph = 0; ph < 3; ++ph) {
        const E3 qs = ((q.lam[ph] * q.invB[ph]) * drive) * (-a);
        const int c = compOfPhase(ph);
        src[c] = src[c] + (qs * f.rho[ph]) * invV;
        if (ph == kO) src[kCG] = src[kCG] + ((q.Rs * qs) * f.rho[kG]) * invV;
    }
}

}  // namespace omodels
