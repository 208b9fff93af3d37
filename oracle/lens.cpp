/*
 * lens.cpp -- CPU restatement (the oracle) of opm-models' element-centred
 * finite-volume path for the two-phase immiscible lens problem
 * (tests/lens_immiscible_ecfv_ad.hh, tests/problems/lensproblem.hh):
 * FvBaseLinearizer::linearize_/linearizeElement_ (fvbaselinearizer.hh:510-626),
 * FvBaseAdLocalLinearizer::linearize (fvbaseadlocallinearizer.hh:173-198,271-292),
 * FvBaseLocalResidual::eval (fvbaselocalresidual.hh:158-191,299-361,419-626),
 * ImmiscibleIntensiveQuantities::update (immiscibleintensivequantities.hh:93-151),
 * ImmiscibleLocalResidual (immisciblelocalresidual.hh:75-155),
 * DarcyExtensiveQuantities (darcyfluxmodule.hh:183-525),
 * FvBaseGradientCalculator (fvbasegradientcalculator.hh:201-300),
 * ImmiscibleBoundaryRateVector::setFreeFlow (immiscibleboundaryratevector.hh:97-159),
 * and the Newton / time loop of newtonmethod.hh:228-465, fvbaseproblem.hh:512-606,
 * simulator.hh:481-512,697-810.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Choices for code outside opm-models (opm-common, dune; absent here, "parity
 * unpinned" against them):
 *  - SimpleH2O liquid: rho 1000 kg/m^3, mu 1e-3 Pa s; DNAPL: rho 1460, mu 5.7e-4.
 *  - EffToAbsLaw: Swe = (Sw - Swr) / (1 - Swr - Snr).
 *  - RegularizedVanGenuchten (m = 1 - 1/n):
 *      pc(Swe)  = (Swe^(-1/m) - 1)^(1/n) / alpha           for 0.01 <= Swe <= 0.99
 *               = pc(0.01) + pc'(0.01) (Swe - 0.01)          for Swe < 0.01
 *               = cubic Hermite from (0.99, pc(0.99), pc'(0.99))
 *                 to (1, 0, m1), m1 = -2 pc(0.99) / 0.01      for 0.99 < Swe < 1
 *               = m1 (Swe - 1)                               for Swe >= 1
 *      krw(Swe) = sqrt(Swe) (1 - (1 - Swe^(1/m))^m)^2, 0 at Swe <= 0, 1 at Swe >= 1
 *      krn(Swe) = (1 - Swe)^(1/3) (1 - Swe^(1/m))^(2m), 1 at Swe <= 0, 0 at Swe >= 1
 *  - harmonicMean(x, y) = 0 if x y <= 0, else 2 x y / (y + x).
 *  - YaspGrid geometry: cell centres (i + 1/2) h, face centres on the cell faces.
 *  - DenseAd pow / sqrt: value by std::pow / std::sqrt, derivative e x^(e-1) x'
 *    and x' / (2 sqrt(x)).
 */
#include "oracle.h"
#include "ad.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <stdexcept>
#include <vector>

namespace {

using namespace oracle_ad;
typedef Ev<2> E2;   /* derivatives w.r.t. (p_w, S_w) of the focus element */

enum { WP = 0, NP = 1 };   /* phases: wetting (water), non-wetting (DNAPL) */

/* ---------------- material law ---------------- */
struct Mat {
    double swr, snr, alpha, n, m;
    double pcLow, slopeLow, pcHigh, slopeHigh, m1;
};

double vgPc(const Mat& p, double swe) { return std::pow(std::pow(swe, -1.0 / p.m) - 1.0, 1.0 / p.n) / p.alpha; }
double vgDpc(const Mat& p, double swe)
{
    return -1.0 / (p.alpha * p.n * p.m) * std::pow(std::pow(swe, -1.0 / p.m) - 1.0, 1.0 / p.n - 1.0) *
           std::pow(swe, -1.0 / p.m - 1.0);
}

Mat makeMat(double swr, double snr, double alpha, double n)
{
    Mat p;
    p.swr = swr; p.snr = snr; p.alpha = alpha; p.n = n; p.m = 1.0 - 1.0 / n;
    p.pcLow = vgPc(p, 0.01);
    p.slopeLow = vgDpc(p, 0.01);
    p.pcHigh = vgPc(p, 0.99);
    p.slopeHigh = vgDpc(p, 0.99);
    p.m1 = (0.0 - p.pcHigh) / (1.0 - 0.99) * 2.0;
    return p;
}

template <class T> T sweOf(const Mat& p, const T& sw) { return (sw - p.swr) / (1.0 - p.swr - p.snr); }

/* plain-number and AD versions share one body through these helpers */
inline double tpow(double a, double e) { return std::pow(a, e); }
inline E2 tpow(const E2& a, double e) { return ev_pow(a, e); }
inline double tsqrt(double a) { return std::sqrt(a); }
inline E2 tsqrt(const E2& a) { return ev_sqrt(a); }
inline double vof(double a) { return a; }
inline double vof(const E2& a) { return a.v; }

template <class T>
T pcnw(const Mat& p, const T& swe)
{
    if (vof(swe) < 0.01) return T(p.pcLow) + (swe - 0.01) * p.slopeLow;
    if (vof(swe) > 0.99) {
        if (vof(swe) < 1.0) {
            const double h = 1.0 - 0.99;
            const T t = (swe - 0.99) / h;
            const T t2 = t * t, t3 = t2 * t;
            const T h00 = t3 * 2.0 - t2 * 3.0 + 1.0;
            const T h10 = t3 - t2 * 2.0 + t;
            const T h01 = t2 * 3.0 - t3 * 2.0;
            const T h11 = t3 - t2;
            return h00 * p.pcHigh + h10 * (h * p.slopeHigh) + h01 * 0.0 + h11 * (h * p.m1);
        }
        return (swe - 1.0) * p.m1;
    }
    return tpow(tpow(swe, -1.0 / p.m) - 1.0, 1.0 / p.n) / p.alpha;
}

template <class T>
T krw(const Mat& p, const T& swe)
{
    if (vof(swe) <= 0.0) return T(0.0);
    if (vof(swe) >= 1.0) return T(1.0);
    const T r = 1.0 - tpow(1.0 - tpow(swe, 1.0 / p.m), p.m);
    return tsqrt(swe) * r * r;
}

template <class T>
T krn(const Mat& p, const T& swe)
{
    if (vof(swe) <= 0.0) return T(1.0);
    if (vof(swe) >= 1.0) return T(0.0);
    return tpow(1.0 - swe, 1.0 / 3.0) * tpow(1.0 - tpow(swe, 1.0 / p.m), 2.0 * p.m);
}

/* ---------------- problem ---------------- */
struct Lens {
    ol_problem P;
    int nx, ny;
    long N;
    double dx, dy, V;
    Mat lens, outer;
    std::vector<uint8_t> inLens;   /* per cell */
    std::vector<int32_t> rowPtr, colIdx, diag;

    explicit Lens(const ol_problem* p) : P(*p)
    {
        nx = p->nx; ny = p->ny; N = (long)nx * ny;
        if (nx < 1 || ny < 1) throw std::invalid_argument("empty grid");
        dx = p->lx / nx; dy = p->ly / ny; V = dx * dy;
        lens = makeMat(p->swr_lens, p->snr_lens, p->alpha_lens, p->n_lens);
        outer = makeMat(p->swr_outer, p->snr_outer, p->alpha_outer, p->n_outer);
        inLens.resize(N);
        for (long I = 0; I < N; ++I) {
            const double x = cx(I), y = cy(I);
            inLens[I] = isInLens(x, y);
        }
        /* pattern (fvbaselinearizer.hh:382-420): sorted columns of the 5-point stencil */
        rowPtr.assign(N + 1, 0);
        for (long I = 0; I < N; ++I) {
            const int i = I % nx, j = I / nx;
            rowPtr[I + 1] = rowPtr[I] + 1 + (i > 0) + (i + 1 < nx) + (j > 0) + (j + 1 < ny);
        }
        colIdx.resize(rowPtr[N]);
        diag.resize(N);
        for (long I = 0; I < N; ++I) {
            const int i = I % nx, j = I / nx;
            int k = rowPtr[I];
            if (j > 0) colIdx[k++] = (int32_t)(I - nx);
            if (i > 0) colIdx[k++] = (int32_t)(I - 1);
            diag[I] = k;
            colIdx[k++] = (int32_t)I;
            if (i + 1 < nx) colIdx[k++] = (int32_t)(I + 1);
            if (j + 1 < ny) colIdx[k++] = (int32_t)(I + nx);
        }
    }
    double cx(long I) const { return (I % nx + 0.5) * dx; }
    double cy(long I) const { return (I / nx + 0.5) * dy; }
    /* lensproblem.hh:590-606 */
    bool isInLens(double x, double y) const
    {
        if (x < P.lens_x0 - P.eps || x > P.lens_x1 + P.eps) return false;
        if (y < P.lens_y0 - P.eps || y > P.lens_y1 + P.eps) return false;
        return true;
    }
    const Mat& mat(long I) const { return inLens[I] ? lens : outer; }
    double K(long I) const { return inLens[I] ? P.k_lens : P.k_outer; }
    long blockOf(long row, long col) const
    {
        for (int k = rowPtr[row]; k < rowPtr[row + 1]; ++k)
            if (colIdx[k] == col) return k;
        throw std::logic_error("block not in pattern");
    }
};

/* ImmiscibleIntensiveQuantities::update (immiscibleintensivequantities.hh:93-151) */
template <class T>
struct IQ {
    T S[2], p[2], mob[2], rho[2];
};

template <class T>
void computeIQ(const Lens& L, long I, const double* pv, IQ<T>& q);

template <>
void computeIQ<E2>(const Lens& L, long I, const double* pv, IQ<E2>& q)
{
    const Mat& m = L.mat(I);
    const E2 sw = E2::var(pv[2 * I + 1], 1);
    const E2 p0 = E2::var(pv[2 * I + 0], 0);
    q.S[WP] = sw;
    q.S[NP] = 1.0 - sw;
    const E2 swe = sweOf(m, sw);
    const E2 pc = pcnw(m, swe);
    q.p[WP] = p0 + (E2(0.0) - E2(0.0));
    q.p[NP] = p0 + (pc - E2(0.0));
    const E2 kw = krw(m, swe), kn = krn(m, swe);
    q.rho[WP] = E2(L.P.rho_w);
    q.rho[NP] = E2(L.P.rho_n);
    q.mob[WP] = kw / E2(L.P.mu_w);
    q.mob[NP] = kn / E2(L.P.mu_n);
}

template <>
void computeIQ<double>(const Lens& L, long I, const double* pv, IQ<double>& q)
{
    IQ<E2> e;
    computeIQ<E2>(L, I, pv, e);
    for (int a = 0; a < 2; ++a) {
        q.S[a] = e.S[a].v; q.p[a] = e.p[a].v; q.mob[a] = e.mob[a].v; q.rho[a] = e.rho[a].v;
    }
}

/* ImmiscibleLocalResidual::computeStorage: phi S_a rho_a (immisciblelocalresidual.hh:75-103) */
template <class T>
void storageOf(const Lens& L, const IQ<T>& q, T st[2])
{
    for (int a = 0; a < 2; ++a) st[a] = (q.S[a] * L.P.porosity) * q.rho[a];
}

double harmonicMean(double x, double y) { return x * y <= 0.0 ? 0.0 : (2.0 * x * y) / (y + x); }

/* Face of element I toward neighbour J: the advective mass flux per unit face
 * area, darcyfluxmodule.hh:183-313 (gradients, gravity correction, upwinding)
 * + :439-456 (filter velocity, volume flux) + immisciblelocalresidual.hh:129-155. */
void interiorFlux(const Lens& L, long I, long J, const IQ<E2>& in, const IQ<double>& ex, const double posI[2],
                  const double posJ[2], const double posF[2], const double nrm[2], E2 flux[2])
{
    const double g[2] = {0.0, L.P.gravity};
    const double Kf = harmonicMean(L.K(I), L.K(J));
    for (int a = 0; a < 2; ++a) {
        /* two-point gradient: (p_J - p_I) d / |d|^2 (fvbasegradientcalculator.hh:201-250) */
        const E2 dy = E2(ex.p[a]) - in.p[a];
        double dist2 = 0.0;
        for (int d = 0; d < 2; ++d) { const double t = posJ[d] - posI[d]; dist2 += t * t; }
        E2 grad[2];
        for (int d = 0; d < 2; ++d) grad[d] = dy * ((posJ[d] - posI[d]) / dist2);
        /* hydrostatic correction (darcyfluxmodule.hh:218-283) */
        double dIn[2], dEx[2], dTot[2];
        for (int d = 0; d < 2; ++d) { dIn[d] = posI[d] - posF[d]; dEx[d] = posJ[d] - posF[d]; dTot[d] = posJ[d] - posI[d]; }
        double absTot2 = 0.0;
        for (int d = 0; d < 2; ++d) absTot2 += dTot[d] * dTot[d];
        double gIn = 0.0, gEx = 0.0;
        for (int d = 0; d < 2; ++d) { gIn += g[d] * dIn[d]; gEx += g[d] * dEx[d]; }
        const E2 pStatIn = -in.rho[a] * gIn;
        const E2 pStatEx = E2(-ex.rho[a] * gEx);
        const E2 fac = (pStatEx - pStatIn) / absTot2;
        for (int d = 0; d < 2; ++d) grad[d] = grad[d] + E2(dTot[d]) * fac;
        /* upstream by the sign of grad . n (ties: interior) */
        E2 tn(0.0);
        for (int d = 0; d < 2; ++d) tn = tn + grad[d] * nrm[d];
        const bool upEx = tn.v > 0.0;
        const E2 mob = upEx ? E2(ex.mob[a]) : in.mob[a];
        const E2 rho = upEx ? E2(ex.rho[a]) : in.rho[a];
        /* filter velocity v = -mob K grad, volume flux v . n */
        E2 vf(0.0);
        for (int d = 0; d < 2; ++d) {
            E2 v = grad[d] * Kf;           /* K is Kf * identity */
            v = v * (-mob);
            vf = vf + v * nrm[d];
        }
        flux[a] = E2(0.0) + vf * rho;
    }
}

/* Boundary face of element I: lensproblem.hh:469-537 with
 * ImmiscibleBoundaryRateVector::setFreeFlow (immiscibleboundaryratevector.hh:97-159)
 * and DarcyExtensiveQuantities::calculateBoundaryGradients_ (darcyfluxmodule.hh:331-437). */
void boundaryRate(const Lens& L, long I, const IQ<E2>& in, const double posI[2], const double posF[2],
                  const double nrm[2], E2 values[2])
{
    const ol_problem& P = L.P;
    values[0] = E2(0.0); values[1] = E2(0.0);
    const bool left = posF[0] < 0.0 + P.eps, right = posF[0] > P.lx - P.eps;
    const bool upper = posF[1] > P.ly - P.eps;
    if (left || right) {
        /* free flow: water-saturated, hydrostatic (left scaled by 1 + 1.5/height) */
        const Mat& m = L.isInLens(posF[0], posF[1]) ? L.lens : L.outer;
        const double depth = P.ly - posF[1];
        double pw;
        if (left) {
            const double alpha = 1.0 + 1.5 / P.ly;
            pw = P.p_atm - alpha * P.rho_w * P.gravity * depth;
        } else {
            pw = P.p_atm - P.rho_w * P.gravity * depth;
        }
        const double sweB = sweOf(m, 1.0);
        const double pcB = pcnw(m, sweB);
        const double pB[2] = {pw, pw + pcB - 0.0};
        const double rhoB[2] = {P.rho_w, P.rho_n};
        const double muB[2] = {P.mu_w, P.mu_n};
        const double krB[2] = {krw(m, sweB), krn(m, sweB)};
        const double g[2] = {0.0, P.gravity};
        const double Kc = L.K(I);
        for (int a = 0; a < 2; ++a) {
            const E2 dy = E2(pB[a]) - in.p[a];
            double dist2 = 0.0;
            for (int d = 0; d < 2; ++d) { const double t = posF[d] - posI[d]; dist2 += t * t; }
            E2 grad[2];
            for (int d = 0; d < 2; ++d) grad[d] = dy * ((posF[d] - posI[d]) / dist2);
            double dIn[2];
            for (int d = 0; d < 2; ++d) dIn[d] = posI[d] - posF[d];
            double abs2 = 0.0, gTd = 0.0;
            for (int d = 0; d < 2; ++d) { abs2 += dIn[d] * dIn[d]; gTd += g[d] * dIn[d]; }
            const E2 pStatIn = in.rho[a] * (-gTd);
            const E2 fac = pStatIn / abs2;
            for (int d = 0; d < 2; ++d) grad[d] = grad[d] + E2(dIn[d]) * fac;
            E2 tn(0.0);
            for (int d = 0; d < 2; ++d) tn = tn + grad[d] * nrm[d];
            const E2 mob = tn.v > 0.0 ? E2(krB[a] / muB[a]) : in.mob[a];
            E2 vf(0.0);
            for (int d = 0; d < 2; ++d) {
                E2 v = grad[d] * Kc;
                v = v * (-mob);
                vf = vf + v * nrm[d];
            }
            const E2 rho = pB[a] > in.p[a].v ? E2(rhoB[a]) : in.rho[a];
            values[a] = values[a] + vf * rho;
        }
    } else if (upper) {
        const double width = P.lx;
        const double lambda = (P.lx - posF[0]) / width;
        if (0.5 < lambda && lambda < 2.0 / 3.0) {
            values[NP] = E2(-P.inflow_rate);   /* kg / (m^2 s), into the domain */
        }
    }
    /* everything else: no flow */
}

/* Element residual of cell I and its derivatives w.r.t. the cell's own
 * primaries (FvBaseLocalResidual::eval, fvbaselocalresidual.hh:158-191),
 * scattered by FvBaseLinearizer::linearizeElement_ (fvbaselinearizer.hh:597-626). */
void linearizeElement(const Lens& L, long I, const double* pv, const double* st1, double dt, double* res,
                      double* jac)
{
    const int nx = L.nx, ny = L.ny;
    const int i = I % nx, j = I / nx;
    IQ<E2> in;
    computeIQ<E2>(L, I, pv, in);
    const double posI[2] = {L.cx(I), L.cy(I)};
    /* local dofs: 0 = element, then neighbours in intersection order */
    E2 r[5][2];
    long dofGlobal[5];
    int ndof = 1;
    dofGlobal[0] = I;
    for (int k = 0; k < 5; ++k) { r[k][0] = E2(0.0); r[k][1] = E2(0.0); }
    /* YaspGrid intersection order: -x, +x, -y, +y */
    const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
    bool boundaryFace[4];
    for (int f = 0; f < 4; ++f) {
        const int ni = i + di[f], nj = j + dj[f];
        const double nrm[2] = {(double)di[f], (double)dj[f]};
        const double posF[2] = {posI[0] + 0.5 * L.dx * di[f], posI[1] + 0.5 * L.dy * dj[f]};
        boundaryFace[f] = !(ni >= 0 && ni < nx && nj >= 0 && nj < ny);
        if (boundaryFace[f]) continue;
        const long J = (long)ni + (long)nx * nj;
        IQ<double> ex;
        computeIQ<double>(L, J, pv, ex);
        const double posJ[2] = {L.cx(J), L.cy(J)};
        E2 flux[2];
        interiorFlux(L, I, J, in, ex, posI, posJ, posF, nrm, flux);
        const double area = 1.0 * (di[f] != 0 ? L.dy : L.dx);
        const int k = ndof++;
        dofGlobal[k] = J;
        for (int e = 0; e < 2; ++e) {
            flux[e] = flux[e] * area;
            r[0][e] = r[0][e] + flux[e];
            r[k][e] = r[k][e] - flux[e];
        }
    }
    /* storage with the cache of the previous time level (fvbaselocalresidual.hh:517-578) */
    E2 st[2];
    storageOf(L, in, st);
    const double fac = L.V / dt;
    for (int e = 0; e < 2; ++e) {
        E2 t = st[e] - st1[2 * I + e];
        t = t * fac;
        r[0][e] = r[0][e] + t;
        r[0][e] = r[0][e] - 0.0 * L.V;   /* source term: zero for this problem */
    }
    /* boundary segments (fvbaselocalresidual.hh:419-484) */
    for (int f = 0; f < 4; ++f) {
        if (!boundaryFace[f]) continue;
        const double nrm[2] = {(double)di[f], (double)dj[f]};
        const double posF[2] = {posI[0] + 0.5 * L.dx * di[f], posI[1] + 0.5 * L.dy * dj[f]};
        E2 values[2];
        boundaryRate(L, I, in, posI, posF, nrm, values);
        const double area = di[f] != 0 ? L.dy : L.dx;
        for (int e = 0; e < 2; ++e) r[0][e] = r[0][e] + values[e] * (area * 1.0);
    }
    /* volume-specific residual */
    for (int k = 0; k < ndof; ++k)
        for (int e = 0; e < 2; ++e) r[k][e] = r[k][e] / L.V;
    for (int e = 0; e < 2; ++e) res[2 * I + e] = r[0][e].v;
    for (int k = 0; k < ndof; ++k) {
        const long b = L.blockOf(dofGlobal[k], I);
        for (int e = 0; e < 2; ++e)
            for (int q = 0; q < 2; ++q) jac[4 * b + 2 * e + q] = r[k][e].d[q];
    }
}

void storageAll(const Lens& L, const double* pv, double* st)
{
    for (long I = 0; I < L.N; ++I) {
        IQ<double> q;
        computeIQ<double>(L, I, pv, q);
        storageOf(L, q, st + 2 * I);
    }
}

void linearizeAll(const Lens& L, const double* pv, const double* st1, double dt, double* res, double* jac)
{
    for (long I = 0; I < L.N; ++I) linearizeElement(L, I, pv, st1, dt, res, jac);
}

}  // namespace

extern "C" {

/* lensproblem.hh:227-627 with the defaults of :254-262,301-320,370-373 */
void ol_default_problem(ol_problem* p)
{
    p->nx = 48; p->ny = 32;
    p->lx = 6.0; p->ly = 4.0;
    p->lens_x0 = 1.0; p->lens_y0 = 2.0; p->lens_x1 = 4.0; p->lens_y1 = 3.0;
    p->k_lens = 9.05e-12; p->k_outer = 4.6e-10;
    p->porosity = 0.4;
    p->swr_lens = 0.18; p->snr_lens = 0.0; p->alpha_lens = 0.00045; p->n_lens = 7.3;
    p->swr_outer = 0.05; p->snr_outer = 0.0; p->alpha_outer = 0.0037; p->n_outer = 4.7;
    p->rho_w = 1000.0; p->rho_n = 1460.0; p->mu_w = 1.0e-3; p->mu_n = 5.7e-4;
    p->gravity = -9.81;
    p->inflow_rate = 0.04;
    p->p_atm = 1e5;
    p->eps = 3e-6;
}

int ol_num_cells(const ol_problem* p, int64_t* n, int64_t* nnzb)
{
    try {
        Lens L(p);
        *n = L.N; *nnzb = L.rowPtr[L.N];
        return 0;
    } catch (...) { return -1; }
}

int ol_pattern(const ol_problem* p, int32_t* row_ptr, int32_t* col_idx, int32_t* diag)
{
    try {
        Lens L(p);
        std::copy(L.rowPtr.begin(), L.rowPtr.end(), row_ptr);
        std::copy(L.colIdx.begin(), L.colIdx.end(), col_idx);
        if (diag) std::copy(L.diag.begin(), L.diag.end(), diag);
        return 0;
    } catch (...) { return -1; }
}

/* LensProblem::initial (lensproblem.hh:550-582): water-saturated, hydrostatic */
int ol_initial(const ol_problem* p, double* pv)
{
    try {
        Lens L(p);
        for (long I = 0; I < L.N; ++I) {
            const double depth = p->ly - L.cy(I);
            pv[2 * I + 0] = p->p_atm - p->rho_w * p->gravity * depth;
            pv[2 * I + 1] = 1.0;
        }
        return 0;
    } catch (...) { return -1; }
}

int ol_storage(const ol_problem* p, const double* pv, double* st)
{
    try { Lens L(p); storageAll(L, pv, st); return 0; } catch (...) { return -1; }
}

int ol_linearize(const ol_problem* p, const double* pv, const double* st1, double dt, double* res, double* jac)
{
    try {
        Lens L(p);
        std::fill(jac, jac + 4 * (long)L.rowPtr[L.N], 0.0);
        linearizeAll(L, pv, st1, dt, res, jac);
        return 0;
    } catch (const std::exception& e) {
        fprintf(stderr, "ol_linearize: %s\n", e.what());
        return -1;
    }
}

/* material law of one region (lens != 0: lens parameters) at S_w:
 * out = {pcnw, dpcnw/dSw, krw, dkrw/dSw, krn, dkrn/dSw} */
int ol_material(const ol_problem* p, int lens, double sw, double* out)
{
    const Mat m = lens ? makeMat(p->swr_lens, p->snr_lens, p->alpha_lens, p->n_lens)
                       : makeMat(p->swr_outer, p->snr_outer, p->alpha_outer, p->n_outer);
    const E2 s = E2::var(sw, 1);
    const E2 swe = sweOf(m, s);
    const E2 pc = pcnw(m, swe), kw = krw(m, swe), kn = krn(m, swe);
    out[0] = pc.v; out[1] = pc.d[1]; out[2] = kw.v; out[3] = kw.d[1]; out[4] = kn.v; out[5] = kn.d[1];
    return 0;
}

/* Simulator::run with end time (simulator.hh:481-512,697-810), timeIntegration
 * with halving (fvbaseproblem.hh:512-557), nextTimeStepSize (:591-606),
 * NewtonMethod::apply (newtonmethod.hh:228-438) with x <- x - dx (:723-731). */
int ol_simulate(const ol_problem* p, double* pv, double t_end, double dt0, const ob_params* prm, int max_steps,
                double* t_hist, double* dt_hist, int32_t* n_steps, ob_iter_log* log, int log_cap,
                int32_t* log_len)
{
    try {
        Lens L(p);
        const long N = L.N;
        const int32_t* rp = L.rowPtr.data();
        const int32_t* ci = L.colIdx.data();
        const long nnzb = rp[N];
        std::vector<double> sol1(pv, pv + 2 * N), st1(2 * N), res(2 * N), jac(4 * nnzb), upd(2 * N);
        int nlog = 0, steps = 0;
        double t = 0.0;
        double dt = std::min(dt0, t_end);
        auto finished = [&]() {
            const double eps = std::max(std::abs(t), dt) * std::numeric_limits<double>::epsilon() * 1e3;
            return t * (1.0 + eps) >= t_end;
        };
        while (!finished()) {
            if (steps >= max_steps) break;
            bool stepOk = false;
            int numIter = 0;
            for (int attempt = 0; attempt < prm->max_dt_divisions; ++attempt) {
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
                    if (numIter == 0) storageAll(L, pv, st1.data());
                    std::fill(jac.begin(), jac.end(), 0.0);
                    linearizeAll(L, pv, st1.data(), dt, res.data(), jac.data());
                    lastError = error;
                    error = 0.0;
                    for (long k = 0; k < 2 * N; ++k) error = std::max(std::abs(res[k]), error);
                    if (error > prm->newton_max_error) { failed = true; break; }
                    if (!proceed()) {
                        if (nlog < log_cap && log) log[nlog++] = {steps, attempt, numIter, 0, 0, 1, error, dt};
                        ++numIter;
                        break;
                    }
                    int32_t it = 0;
                    const int rc = ob_bicgstab((int)N, 2, rp, ci, jac.data(), res.data(), upd.data(), prm, &it);
                    if (rc != 0) {
                        if (nlog < log_cap && log) log[nlog++] = {steps, attempt, numIter, it, 0, 0, error, dt};
                        failed = true;
                        break;
                    }
                    bool finite = true;
                    for (long k = 0; k < 2 * N; ++k) {
                        if (!std::isfinite(upd[k])) { finite = false; break; }
                        pv[k] = pv[k] - upd[k];
                    }
                    if (!finite) { failed = true; break; }
                    if (nlog < log_cap && log) log[nlog++] = {steps, attempt, numIter, it, 0, 1, error, dt};
                    ++numIter;
                }
                if (!failed && converged()) { stepOk = true; break; }
                numIter = 2 * prm->newton_target_iter;
                std::copy(sol1.begin(), sol1.end(), pv);
                const double ndt = dt / 2.0;
                if (ndt < prm->min_dt) break;
                dt = ndt;
            }
            if (!stepOk) {
                if (n_steps) *n_steps = steps;
                if (log_len) *log_len = nlog;
                return 3;
            }
            std::copy(pv, pv + 2 * N, sol1.begin());
            t += dt;
            if (t_hist) t_hist[steps] = t;
            if (dt_hist) dt_hist[steps] = dt;
            ++steps;
            /* nextTimeStepSize, clipped to the end time */
            const int tgt = prm->newton_target_iter;
            double sug;
            if (numIter > tgt) sug = dt / (1.0 + double(numIter - tgt) / tgt);
            else sug = dt * (1.0 + double(tgt - numIter) / tgt / 1.2);
            double dtNext = std::min(prm->max_dt, sug);
            const double remain = std::max(0.0, t_end - t);
            if (dtNext < remain && remain < dtNext * 2) dtNext = remain / 2 * 1.01;
            dt = std::min(remain, dtNext);
        }
        if (n_steps) *n_steps = steps;
        if (log_len) *log_len = nlog;
        return 0;
    } catch (const std::exception& e) {
        fprintf(stderr, "ol_simulate: %s\n", e.what());
        return -1;
    }
}

}  // extern "C"
