/*
 * ad.h -- forward-mode AD scalar of the oracle (test infrastructure only).
 */
#ifndef OMODELS_ORACLE_AD_H
#define OMODELS_ORACLE_AD_H

#include <cmath>

namespace oracle_ad {

/* ------------------------------------------------------------------ */
/* Forward-mode AD scalar (restates DenseAd::Evaluation semantics as used by
 * fvbaseadlocallinearizer.hh:62-72 and blackoilprimaryvariables.hh:217-233). */
template <int N>
struct Ev {
    double v;
    double d[N];
    Ev() : v(0.0) { for (int i = 0; i < N; ++i) d[i] = 0.0; }
    Ev(double x) : v(x) { for (int i = 0; i < N; ++i) d[i] = 0.0; }
    static Ev var(double x, int idx) { Ev r(x); r.d[idx] = 1.0; return r; }
};
template <int N> inline Ev<N> operator+(const Ev<N>& a, const Ev<N>& b)
{ Ev<N> r; r.v = a.v + b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> inline Ev<N> operator-(const Ev<N>& a, const Ev<N>& b)
{ Ev<N> r; r.v = a.v - b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> inline Ev<N> operator*(const Ev<N>& a, const Ev<N>& b)
{ Ev<N> r; r.v = a.v * b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int N> inline Ev<N> operator/(const Ev<N>& a, const Ev<N>& b)
{
    Ev<N> r; const double inv = 1.0 / b.v; r.v = a.v * inv;
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
    return r;
}
template <int N> inline Ev<N> operator+(const Ev<N>& a, double s)
{ Ev<N> r = a; r.v = a.v + s; return r; }
template <int N> inline Ev<N> operator-(const Ev<N>& a, double s)
{ Ev<N> r = a; r.v = a.v - s; return r; }
template <int N> inline Ev<N> operator-(double s, const Ev<N>& a)
{ Ev<N> r; r.v = s - a.v; for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r; }
template <int N> inline Ev<N> operator*(const Ev<N>& a, double s)
{ Ev<N> r; r.v = a.v * s; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * s; return r; }
template <int N> inline Ev<N> operator*(double s, const Ev<N>& a) { return a * s; }
template <int N> inline Ev<N> ev_min(const Ev<N>& a, double c) { return a.v <= c ? a : Ev<N>(c); }
template <int N> inline Ev<N> ev_max(const Ev<N>& a, double c) { return a.v >= c ? a : Ev<N>(c); }

/* division by a plain number: value and derivatives divided (Evaluation / Scalar) */
template <int N> inline Ev<N> operator/(const Ev<N>& a, double s)
{ Ev<N> r; r.v = a.v / s; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / s; return r; }
template <int N> inline Ev<N> operator-(const Ev<N>& a)
{ Ev<N> r; r.v = -a.v; for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r; }
/* pow(x, e) = x^e, derivative e x^(e-1) x' ; sqrt(x), derivative x' / (2 sqrt(x)) */
template <int N> inline Ev<N> ev_pow(const Ev<N>& a, double e)
{
    Ev<N> r; r.v = std::pow(a.v, e);
    const double df = e * std::pow(a.v, e - 1.0);
    for (int i = 0; i < N; ++i) r.d[i] = df * a.d[i];
    return r;
}
template <int N> inline Ev<N> ev_sqrt(const Ev<N>& a)
{
    Ev<N> r; r.v = std::sqrt(a.v);
    const double df = 0.5 / r.v;
    for (int i = 0; i < N; ++i) r.d[i] = df * a.d[i];
    return r;
}

}  // namespace oracle_ad

#endif
