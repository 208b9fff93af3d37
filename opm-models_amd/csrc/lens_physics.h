// lens_physics.h -- device physics of the two-phase immiscible element path
// (the lens problem, tests/problems/lensproblem.hh): AD scalar with the two
// derivatives of an element's own primaries (p_w, S_w), regularized Van
// Genuchten through EffToAbsLaw, intensive quantities, storage, the Darcy
// flux of an interior face and the boundary rates.  Every expression keeps
// the operation order of the restatement it is tested against, so that the
// residual and Jacobian agree to rounding of the math library's pow/sqrt.
#pragma once

#include <hip/hip_runtime.h>

namespace omodels {

// DenseAd::Evaluation<double, 2>
struct E2 {
    double v, d0, d1;
};
__device__ __forceinline__ E2 e2(double x) { return E2{x, 0.0, 0.0}; }
__device__ __forceinline__ E2 operator+(E2 a, E2 b) { return E2{a.v + b.v, a.d0 + b.d0, a.d1 + b.d1}; }
__device__ __forceinline__ E2 operator-(E2 a, E2 b) { return E2{a.v - b.v, a.d0 - b.d0, a.d1 - b.d1}; }
__device__ __forceinline__ E2 operator*(E2 a, E2 b)
{
    return E2{a.v * b.v, a.d0 * b.v + a.v * b.d0, a.d1 * b.v + a.v * b.d1};
}
__device__ __forceinline__ E2 operator/(E2 a, E2 b)
{
    const double inv = 1.0 / b.v;
    const double v = a.v * inv;
    return E2{v, (a.d0 - v * b.d0) * inv, (a.d1 - v * b.d1) * inv};
}
__device__ __forceinline__ E2 operator+(E2 a, double s) { return E2{a.v + s, a.d0, a.d1}; }
__device__ __forceinline__ E2 operator-(E2 a, double s) { return E2{a.v - s, a.d0, a.d1}; }
__device__ __forceinline__ E2 operator-(double s, E2 a) { return E2{s - a.v, -a.d0, -a.d1}; }
__device__ __forceinline__ E2 operator*(E2 a, double s) { return E2{a.v * s, a.d0 * s, a.d1 * s}; }
__device__ __forceinline__ E2 operator/(E2 a, double s) { return E2{a.v / s, a.d0 / s, a.d1 / s}; }
__device__ __forceinline__ E2 operator-(E2 a) { return E2{-a.v, -a.d0, -a.d1}; }
__device__ __forceinline__ E2 e2pow(E2 a, double e)
{
    const double df = e * pow(a.v, e - 1.0);
    return E2{pow(a.v, e), df * a.d0, df * a.d1};
}
__device__ __forceinline__ E2 e2sqrt(E2 a)
{
    const double v = sqrt(a.v);
    const double df = 0.5 / v;
    return E2{v, df * a.d0, df * a.d1};
}

// Regularized Van Genuchten parameters (the finalized values: m = 1 - 1/n,
// pc and its slope at Swe 0.01 and 0.99, the end slope m1 at Swe = 1)
struct LensMat {
    double swr, snr, alpha, n, m;
    double pcLow, slopeLow, pcHigh, slopeHigh, m1;
};

struct LensParams {
    int nx, ny;
    double lx, ly, dx, dy, V;
    double lensX0, lensY0, lensX1, lensY1;
    double kLens, kOuter, porosity;
    LensMat lens, outer;
    double rhoW, rhoN, muW, muN, g, inflow, pAtm, eps;
};

__device__ __forceinline__ E2 lensSwe(const LensMat& p, E2 sw) { return (sw - p.swr) / (1.0 - p.swr - p.snr); }

__device__ __forceinline__ E2 lensPcnw(const LensMat& p, E2 swe)
{
    if (swe.v < 0.01) return e2(p.pcLow) + (swe - 0.01) * p.slopeLow;
    if (swe.v > 0.99) {
        if (swe.v < 1.0) {
            const double h = 1.0 - 0.99;
            const E2 t = (swe - 0.99) / h;
            const E2 t2 = t * t, t3 = t2 * t;
            const E2 h00 = t3 * 2.0 - t2 * 3.0 + 1.0;
            const E2 h10 = t3 - t2 * 2.0 + t;
            const E2 h01 = t2 * 3.0 - t3 * 2.0;
            const E2 h11 = t3 - t2;
            return h00 * p.pcHigh + h10 * (h * p.slopeHigh) + h01 * 0.0 + h11 * (h * p.m1);
        }
        return (swe - 1.0) * p.m1;
    }
    return e2pow(e2pow(swe, -1.0 / p.m) - 1.0, 1.0 / p.n) / p.alpha;
}

__device__ __forceinline__ E2 lensKrw(const LensMat& p, E2 swe)
{
    if (swe.v <= 0.0) return e2(0.0);
    if (swe.v >= 1.0) return e2(1.0);
    const E2 r = 1.0 - e2pow(1.0 - e2pow(swe, 1.0 / p.m), p.m);
    return e2sqrt(swe) * r * r;
}

__device__ __forceinline__ E2 lensKrn(const LensMat& p, E2 swe)
{
    if (swe.v <= 0.0) return e2(1.0);
    if (swe.v >= 1.0) return e2(0.0);
    return e2pow(1.0 - swe, 1.0 / 3.0) * e2pow(1.0 - e2pow(swe, 1.0 / p.m), 2.0 * p.m);
}

// ImmiscibleIntensiveQuantities::update (immiscibleintensivequantities.hh:93-151)
struct LensIQ {
    E2 S[2], p[2], mob[2], rho[2];
};

__device__ __forceinline__ void lensIQ(const LensParams& P, bool inLens, double pw, double sw0, LensIQ& q)
{
    const LensMat& m = inLens ? P.lens : P.outer;
    const E2 sw = E2{sw0, 0.0, 1.0};
    const E2 p0 = E2{pw, 1.0, 0.0};
    q.S[0] = sw;
    q.S[1] = 1.0 - sw;
    const E2 swe = lensSwe(m, sw);
    const E2 pc = lensPcnw(m, swe);
    q.p[0] = p0 + (e2(0.0) - e2(0.0));
    q.p[1] = p0 + (pc - e2(0.0));
    const E2 kw = lensKrw(m, swe), kn = lensKrn(m, swe);
    q.rho[0] = e2(P.rhoW);
    q.rho[1] = e2(P.rhoN);
    q.mob[0] = kw / e2(P.muW);
    q.mob[1] = kn / e2(P.muN);
}

// values only (the neighbour side of a face)
struct LensIQv {
    double p[2], mob[2], rho[2];
};
__device__ __forceinline__ void lensIQv(const LensParams& P, bool inLens, double pw, double sw0, LensIQv& q)
{
    LensIQ e;
    lensIQ(P, inLens, pw, sw0, e);
    for (int a = 0; a < 2; ++a) { q.p[a] = e.p[a].v; q.mob[a] = e.mob[a].v; q.rho[a] = e.rho[a].v; }
}

__device__ __forceinline__ double lensHarmonicMean(double x, double y)
{
    return x * y <= 0.0 ? 0.0 : (2.0 * x * y) / (y + x);
}

// Darcy flux over an interior face of the focus element (darcyfluxmodule.hh:183-313,439-456,
// immisciblelocalresidual.hh:129-155): mass flux per unit face area per phase
__device__ __forceinline__ void lensInteriorFlux(const LensParams& P, double Kf, const LensIQ& in,
                                                 const LensIQv& ex, const double posI[2], const double posJ[2],
                                                 const double posF[2], const double nrm[2], E2 flux[2])
{
    const double g[2] = {0.0, P.g};
    for (int a = 0; a < 2; ++a) {
        const E2 dy = e2(ex.p[a]) - in.p[a];
        double dist2 = 0.0;
        for (int d = 0; d < 2; ++d) { const double t = posJ[d] - posI[d]; dist2 += t * t; }
        E2 grad[2];
        for (int d = 0; d < 2; ++d) grad[d] = dy * ((posJ[d] - posI[d]) / dist2);
        double dIn[2], dEx[2], dTot[2];
        for (int d = 0; d < 2; ++d) { dIn[d] = posI[d] - posF[d]; dEx[d] = posJ[d] - posF[d]; dTot[d] = posJ[d] - posI[d]; }
        double absTot2 = 0.0;
        for (int d = 0; d < 2; ++d) absTot2 += dTot[d] * dTot[d];
        double gIn = 0.0, gEx = 0.0;
        for (int d = 0; d < 2; ++d) { gIn += g[d] * dIn[d]; gEx += g[d] * dEx[d]; }
        const E2 pStatIn = -in.rho[a] * gIn;
        const E2 pStatEx = e2(-ex.rho[a] * gEx);
        const E2 fac = (pStatEx - pStatIn) / absTot2;
        for (int d = 0; d < 2; ++d) grad[d] = grad[d] + e2(dTot[d]) * fac;
        E2 tn = e2(0.0);
        for (int d = 0; d < 2; ++d) tn = tn + grad[d] * nrm[d];
        const bool upEx = tn.v > 0.0;
        const E2 mob = upEx ? e2(ex.mob[a]) : in.mob[a];
        const E2 rho = upEx ? e2(ex.rho[a]) : in.rho[a];
        E2 vf = e2(0.0);
        for (int d = 0; d < 2; ++d) {
            E2 v = grad[d] * Kf;
            v = v * (-mob);
            vf = vf + v * nrm[d];
        }
        flux[a] = e2(0.0) + vf * rho;
    }
}

__device__ __forceinline__ bool lensInBox(const LensParams& P, double x, double y)
{
    if (x < P.lensX0 - P.eps || x > P.lensX1 + P.eps) return false;
    if (y < P.lensY0 - P.eps || y > P.lensY1 + P.eps) return false;
    return true;
}

// LensProblem::boundary (lensproblem.hh:469-537) with setFreeFlow
// (immiscibleboundaryratevector.hh:97-159) and the boundary gradients of
// darcyfluxmodule.hh:331-437: rate per unit face area, positive out of the domain
__device__ __forceinline__ void lensBoundaryRate(const LensParams& P, double Kc, const LensIQ& in,
                                                 const double posI[2], const double posF[2], const double nrm[2],
                                                 E2 values[2])
{
    values[0] = e2(0.0);
    values[1] = e2(0.0);
    const bool left = posF[0] < 0.0 + P.eps, right = posF[0] > P.lx - P.eps;
    const bool upper = posF[1] > P.ly - P.eps;
    if (left || right) {
        const LensMat& m = lensInBox(P, posF[0], posF[1]) ? P.lens : P.outer;
        const double depth = P.ly - posF[1];
        double pw;
        if (left) {
            const double alpha = 1.0 + 1.5 / P.ly;
            pw = P.pAtm - alpha * P.rhoW * P.g * depth;
        } else {
            pw = P.pAtm - P.rhoW * P.g * depth;
        }
        const double sweB = lensSwe(m, e2(1.0)).v;
        const double pcB = lensPcnw(m, e2(sweB)).v;
        const double pB[2] = {pw, pw + pcB - 0.0};
        const double rhoB[2] = {P.rhoW, P.rhoN};
        const double muB[2] = {P.muW, P.muN};
        const double krB[2] = {lensKrw(m, e2(sweB)).v, lensKrn(m, e2(sweB)).v};
        const double g[2] = {0.0, P.g};
        for (int a = 0; a < 2; ++a) {
            const E2 dy = e2(pB[a]) - in.p[a];
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
            for (int d = 0; d < 2; ++d) grad[d] = grad[d] + e2(dIn[d]) * fac;
            E2 tn = e2(0.0);
            for (int d = 0; d < 2; ++d) tn = tn + grad[d] * nrm[d];
            const E2 mob = tn.v > 0.0 ? e2(krB[a] / muB[a]) : in.mob[a];
            E2 vf = e2(0.0);
            for (int d = 0; d < 2; ++d) {
                E2 v = grad[d] * Kc;
                v = v * (-mob);
                vf = vf + v * nrm[d];
            }
            const E2 rho = pB[a] > in.p[a].v ? e2(rhoB[a]) : in.rho[a];
            values[a] = values[a] + vf * rho;
        }
    } else if (upper) {
        const double width = P.lx;
        const double lambda = (P.lx - posF[0]) / width;
        if (0.5 < lambda && lambda < 2.0 / 3.0) values[1] = e2(-P.inflow);
    }
}

}  // namespace omodels
