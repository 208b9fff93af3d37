// problem.cpp -- black-oil problem data on the host: the fluid tables of the
// reference's ReservoirProblem, a deterministic synthetic generator for the
// benchmark configurations, and the structured-grid topology that
// TpfaLinearizer::createMatrix_ builds (tpfalinearizer.hh:414-525).
#include <cmath>
#include <cstring>
#include <algorithm>
#include <limits>

#include "omodels_host.h"

namespace omodels {

// ---------------------------------------------------------------------------
// Fluid of tests/problems/reservoirproblem.hh:223-327 (LiveOil, DryGas,
// ConstantCompressibilityWater; surface densities 786 / 0.97 / 1037).
static const double kPo[9] = {101353, 1.82504e+06, 3.54873e+06, 6.99611e+06, 1.38909e+07,
                              1.73382e+07, 2.07856e+07, 2.76804e+07, 3.45751e+07};
static const double kBo[9] = {1.062, 1.15, 1.207, 1.295, 1.435, 1.5, 1.565, 1.695, 1.827};
static const double kMuo[9] = {0.00104, 0.000975, 0.00091, 0.00083, 0.000695,
                               0.000641, 0.000594, 0.00051, 0.000449};
static const double kRs[9] = {0.178108, 16.1187, 32.0594, 66.0779, 113.276,
                              138.033, 165.64, 226.197, 288.178};
static const double kPg[10] = {101353, 1.82504e+06, 3.54873e+06, 6.99611e+06, 1.38909e+07,
                               1.73382e+07, 2.07856e+07, 2.76804e+07, 3.45751e+07, 6.21542e+07};
static const double kBg[10] = {0.93576, 0.0678972, 0.0352259, 0.0179498, 0.00906194,
                               0.00726527, 0.00606375, 0.00455343, 0.00364386, 0.00216723};
static const double kMug[10] = {8e-06, 9.6e-06, 1.12e-05, 1.4e-05, 1.89e-05,
                                2.08e-05, 2.28e-05, 2.68e-05, 3.09e-05, 4.7e-05};

void defaultFluid(std::vector<double>& oil, int& nOil, std::vector<double>& gas, int& nGas, double c[7])
{
    nOil = 9; nGas = 10;
    oil.assign(4 * 9, 0.0); gas.assign(3 * 10, 0.0);
    for (int i = 0; i < 9; ++i) {
        oil[i] = kPo[i]; oil[9 + i] = kRs[i]; oil[18 + i] = 1.0 / kBo[i]; oil[27 + i] = kMuo[i];
    }
    for (int i = 0; i < 10; ++i) { gas[i] = kPg[i]; gas[10 + i] = 1.0 / kBg[i]; gas[20 + i] = kMug[i]; }
    // oil compressibility of the undersaturated extension: the density slope
    // (1.1200 - 1.1189) / ((5000 - 4000) psi) used for LiveOil tables without
    // undersaturated data, in 1/Pa
    c[0] = (1.1200 - 1.1189) / ((5000.0 - 4000.0) * 6894.76);
    c[1] = 101325.0;          // water reference pressure
    c[2] = 1.450377e-10;      // water compressibility
    c[3] = 9.6e-4;            // water viscosity
    c[4] = 1037.0; c[5] = 786.0; c[6] = 0.97;   // rho_ref water, oil, gas
}

void fillTab(Tab& t, const double* x, const double* y, int n)
{
    if (n < 2 || n > kMaxTab) throw std::invalid_argument("table size must be in [2, 12]");
    t.n = n;
    for (int i = 0; i < kMaxTab; ++i) { t.x[i] = 0.0; t.y[i] = 0.0; t.s[i] = 0.0; }
    for (int i = 0; i < n; ++i) { t.x[i] = x[i]; t.y[i] = y[i]; }
    for (int i = 0; i + 1 < n; ++i) t.s[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
    for (int i = 1; i < n; ++i)
        if (!(x[i] > x[i - 1])) throw std::invalid_argument("table abscissae must increase");
}

BoFluid makeFluid(const omodels_blackoil_problem* p)
{
    BoFluid f;
    std::memset(&f, 0, sizeof(f));
    const int n = p->n_oil; const double* t = p->oil_tab;
    fillTab(f.rsSat, t, t + n, n);
    fillTab(f.pOfRs, t + n, t, n);
    fillTab(f.invBoSat, t, t + 2 * n, n);
    fillTab(f.muoSat, t, t + 3 * n, n);
    const int m = p->n_gas; const double* g = p->gas_tab;
    fillTab(f.invBg, g, g + m, m);
    fillTab(f.mug, g, g + 2 * m, m);
    f.co = p->oil_comp; f.pref = p->water_pref; f.cw = p->water_comp; f.muw = p->water_mu;
    for (int a = 0; a < 3; ++a) f.rho[a] = p->rho_ref[a];
    return f;
}

// host mirror of the device table evaluation (generator only)
static double tabHost(const Tab& t, double x)
{
    double x0 = t.x[0], y0 = t.y[0], s0 = t.s[0];
    for (int i = 1; i + 1 < t.n; ++i) if (x >= t.x[i]) { x0 = t.x[i]; y0 = t.y[i]; s0 = t.s[i]; }
    return y0 + s0 * (x - x0);
}

// ---------------------------------------------------------------------------
// Counter-based RNG (SplitMix64 finalizer): value = f(seed, stream, index).
static inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline double uniform01(uint64_t seed, uint64_t stream, uint64_t idx)
{
    const uint64_t h = mix64(seed ^ mix64(stream * 0x632BE59BD9B4E019ull + idx));
    return ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
static inline double normal01(uint64_t seed, uint64_t stream, uint64_t idx)
{
    const double u1 = uniform01(seed, stream, 2 * idx), u2 = uniform01(seed, stream, 2 * idx + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

void generate(const omodels_blackoil_spec& sp, std::vector<double>& perm, std::vector<double>& poro,
              std::vector<double>& pv, std::vector<uint8_t>& mean, std::vector<int32_t>& srcCell,
              std::vector<int32_t>& srcType, std::vector<double>& srcA, std::vector<double>& srcB)
{
    const long nx = sp.nx, ny = sp.ny, nz = sp.nz, N = nx * ny * nz;
    if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("grid dimensions must be positive");
    const double mD = 9.869233e-16;
    perm.assign(3 * N, 0.0); poro.assign(N, 0.0); pv.assign(3 * N, 0.0); mean.assign(N, 0);
    // channelised field: a seeded sum of Fourier modes, mostly elongated in x
    std::vector<double> gauss;
    double thresh = 0.0;
    if (sp.field == 1) {
        const int nm = 48;
        std::vector<double> kxv(nm), kyv(nm), kzv(nm), ph(nm);
        const double L = sp.corr_len > 0 ? sp.corr_len : 10.0 * sp.dx;
        for (int m = 0; m < nm; ++m) {
            const double ang = 0.35 * (2.0 * uniform01(sp.seed, 11, m) - 1.0);
            const double kk = (0.5 + uniform01(sp.seed, 12, m)) * 6.283185307179586 / L;
            kxv[m] = kk * std::cos(ang) * 0.25; kyv[m] = kk * std::sin(ang) + kk * 0.6;
            kzv[m] = kk * 0.3 * (2.0 * uniform01(sp.seed, 13, m) - 1.0);
            ph[m] = 6.283185307179586 * uniform01(sp.seed, 14, m);
        }
        gauss.resize(N);
        for (long I = 0; I < N; ++I) {
            const long i = I % nx, j = (I / nx) % ny, k = I / (nx * ny);
            const double x = (i + 0.5) * sp.dx, y = (j + 0.5) * sp.dy, z = (k + 0.5) * sp.dz;
            double s = 0.0;
            for (int m = 0; m < nm; ++m) s += std::cos(kxv[m] * x + kyv[m] * y + kzv[m] * z + ph[m]);
            gauss[I] = s * std::sqrt(2.0 / nm);
        }
        std::vector<double> tmp(gauss);
        const double frac = (sp.chan_fraction > 0 && sp.chan_fraction < 1) ? sp.chan_fraction : 0.3;
        const long q = std::min<long>(N - 1, (long)((1.0 - frac) * N));
        std::nth_element(tmp.begin(), tmp.begin() + q, tmp.end());
        thresh = tmp[q];
    }
    for (long I = 0; I < N; ++I) {
        double lnk;
        const double z = normal01(sp.seed, 1, I);
        if (sp.field == 1) {
            const double base = gauss[I] >= thresh ? sp.chan_perm_md : sp.back_perm_md;
            lnk = std::log(base * mD) + sp.perm_sigma * z;
        } else {
            lnk = std::log(sp.perm_mean_md * mD) + sp.perm_sigma * z;
        }
        const double kx = std::exp(lnk);
        perm[I] = kx; perm[N + I] = kx; perm[2 * N + I] = sp.kz_ratio * kx;
        poro[I] = sp.poro_min + (sp.poro_max - sp.poro_min) * uniform01(sp.seed, 2, I);
    }
    // hydrostatic oil-phase equilibrium of saturated oil, top layer at p_top:
    // p_below = p_above + (rho_above + rho_below)/2 * |g| * dz (the discrete
    // balance of the TPFA pressure difference with zero flux)
    std::vector<double> oil; int nOil; std::vector<double> gas; int nGas; double c[7];
    defaultFluid(oil, nOil, gas, nGas, c);
    Tab rsSat, invBo;
    fillTab(rsSat, oil.data(), oil.data() + nOil, nOil);
    fillTab(invBo, oil.data(), oil.data() + 2 * nOil, nOil);
    auto rhoOil = [&](double p) { const double b = tabHost(invBo, p); return b * c[5] + (b * tabHost(rsSat, p)) * c[6]; };
    std::vector<double> pk(nz);
    pk[nz - 1] = sp.p_top;
    for (long k = nz - 2; k >= 0; --k) {
        const double pu = pk[k + 1], ru = rhoOil(pu);
        double p = pu;
        for (int it = 0; it < 50; ++it) p = pu + 0.5 * (ru + rhoOil(p)) * 9.81 * sp.dz;
        pk[k] = p;
    }
    for (long I = 0; I < N; ++I) {
        const long k = I / (nx * ny);
        pv[3 * I] = sp.sw_init;
        pv[3 * I + 1] = pk[k];
        const double f = (sp.rs_frac > 0.0 && sp.rs_frac < 1.0) ? sp.rs_frac : 1.0;
        pv[3 * I + 2] = f * tabHost(rsSat, pk[k]);
        mean[I] = kMeanRs;
    }
    // wells as source terms on full columns
    srcCell.clear(); srcType.clear(); srcA.clear(); srcB.clear();
    auto columnKh = [&](long i, long j) {
        double s = 0.0;
        for (long k = 0; k < nz; ++k) s += perm[i + nx * (j + ny * k)];
        return s;
    };
    auto column = [&](long i, long j, int type, double a) {
        // the channelised field's injector gives each layer the share of the
        // rate its kh has of the column's (the completions' allocation); the
        // log-normal fields split it evenly
        const double khSum = columnKh(i, j);
        for (long k = 0; k < nz; ++k) {
            const long I = i + nx * (j + ny * k);
            srcCell.push_back((int32_t)I); srcType.push_back(type);
            if (type == 0) {
                srcA.push_back(sp.field == 1 ? a * (perm[I] / khSum) : a / nz);
                srcB.push_back(0.0);
            } else {
                const double kh = std::sqrt(perm[I] * perm[N + I]);
                const double re = 0.2 * sp.dx, rw = 0.1;
                srcA.push_back(a * 6.283185307179586 * kh * sp.dz / std::log(re / rw));
                srcB.push_back(sp.prod_bhp);
            }
        }
    };
    if (sp.wells == 0) {
        column(0, 0, 0, sp.inj_rate);
        column(nx - 1, ny - 1, 1, 1.0);
    } else if (sp.wells == 1) {
        // the channelised field's injector sits in the column of the largest
        // kh within a few cells of the centre (a well placed in a channel)
        long ci = nx / 2, cj = ny / 2;
        if (sp.field == 1) {
            const long r = std::max<long>(1, std::min(nx, ny) / 20);
            double best = -1.0;
            for (long j = std::max<long>(0, ny / 2 - r); j <= std::min(ny - 1, ny / 2 + r); ++j)
                for (long i = std::max<long>(0, nx / 2 - r); i <= std::min(nx - 1, nx / 2 + r); ++i) {
                    const double v = columnKh(i, j);
                    if (v > best) { best = v; ci = i; cj = j; }
                }
        }
        column(ci, cj, 0, sp.inj_rate);
        column(0, 0, 1, 1.0); column(nx - 1, 0, 1, 1.0);
        column(0, ny - 1, 1, 1.0); column(nx - 1, ny - 1, 1, 1.0);
    } else {
        const long ii = (long)(uniform01(sp.seed, 21, 0) * nx) % nx;
        const long ij = (long)(uniform01(sp.seed, 21, 1) * ny) % ny;
        long pi = (long)(uniform01(sp.seed, 21, 2) * nx) % nx;
        long pj = (long)(uniform01(sp.seed, 21, 3) * ny) % ny;
        if (pi == ii && pj == ij) { pi = (pi + nx / 2) % nx; pj = (pj + ny / 2) % ny; }
        if (pi == ii && pj == ij) pi = (pi + 1) % nx;
        column(ii, ij, 0, sp.inj_rate);
        if (!(pi == ii && pj == ij)) column(pi, pj, 1, 1.0);
    }
}

// ---------------------------------------------------------------------------
// Topology of a structured grid: sorted columns per row (-z,-y,-x,I,+x,+y,+z
// as present), diagonal block offsets, the offset of A[J][I] for each
// intersection of I in the order -x,+x,-y,+y,-z,+z, per-face
// transmissibilities and the ILU level sets.
static inline long nbOf(const HostTopology& T, long I, int dir)
{
    const long i = I % T.nx, j = (I / T.nx) % T.ny, k = I / ((long)T.nx * T.ny);
    switch (dir) {
    case 0: return i > 0 ? I - 1 : -1;
    case 1: return i + 1 < T.nx ? I + 1 : -1;
    case 2: return j > 0 ? I - T.nx : -1;
    case 3: return j + 1 < T.ny ? I + T.nx : -1;
    case 4: return k > 0 ? I - (long)T.nx * T.ny : -1;
    default: return k + 1 < T.nz ? I + (long)T.nx * T.ny : -1;
    }
}

void buildTopology(const omodels_blackoil_problem* p, HostTopology& T)
{
    T.nx = p->nx; T.ny = p->ny; T.nz = p->nz;
    const long N = (long)T.nx * T.ny * T.nz;
    T.N = N;
    static const int sortedDirs[6] = {4, 2, 0, 1, 3, 5};   // column order around the diagonal
    T.rowPtr.assign(N + 1, 0);
    for (long I = 0; I < N; ++I) {
        int c = 1;
        for (int d = 0; d < 6; ++d) c += nbOf(T, I, d) >= 0;
        T.rowPtr[I + 1] = T.rowPtr[I] + c;
    }
    const long nnzb = T.rowPtr[N];
    if (nnzb >= (long)std::numeric_limits<int32_t>::max() / 9)
        throw std::invalid_argument("grid too large for 32-bit block offsets");
    T.colIdx.resize(nnzb);
    T.diagOff.resize(N);
    // position of each present direction in the sorted row
    std::vector<int8_t> pos(6 * N, -1);
    for (long I = 0; I < N; ++I) {
        int k = T.rowPtr[I];
        for (int s = 0; s < 3; ++s) {
            const long J = nbOf(T, I, sortedDirs[s]);
            if (J >= 0) { pos[6 * I + sortedDirs[s]] = (int8_t)(k - T.rowPtr[I]); T.colIdx[k++] = (int32_t)J; }
        }
        T.diagOff[I] = k;
        T.colIdx[k++] = (int32_t)I;
        for (int s = 3; s < 6; ++s) {
            const long J = nbOf(T, I, sortedDirs[s]);
            if (J >= 0) { pos[6 * I + sortedDirs[s]] = (int8_t)(k - T.rowPtr[I]); T.colIdx[k++] = (int32_t)J; }
        }
    }
    // A[J][I]: I is J's neighbour in the opposite direction
    T.nbOff.assign(6 * N, -1);
    for (long I = 0; I < N; ++I)
        for (int d = 0; d < 6; ++d) {
            const long J = nbOf(T, I, d);
            if (J < 0) continue;
            const int opp = d ^ 1;
            T.nbOff[(long)d * N + I] = T.rowPtr[J] + pos[6 * J + opp];
        }
    // transmissibility of the face between I and I + e_axis:
    // harmonic mean of A K |n.d| / |d|^2 (transfluxmodule.hh:425-456)
    T.trans.assign(3 * N, 0.0);
    const double h[3] = {p->dx, p->dy, p->dz};
    const double area[3] = {p->dy * p->dz, p->dx * p->dz, p->dx * p->dy};
    for (long I = 0; I < N; ++I)
        for (int ax = 0; ax < 3; ++ax) {
            const long J = nbOf(T, I, 2 * ax + 1);
            if (J < 0) continue;
            const double A = area[ax], ndot = 0.5 * h[ax], dist2 = ndot * ndot;
            const double K0 = p->perm[(long)ax * N + I], K1 = p->perm[(long)ax * N + J];
            const double T0 = A * K0 * ndot / dist2;
            const double T1 = A * K1 * ndot / dist2;
            T.trans[(long)ax * N + I] = T0 * T1 / (T0 + T1);
        }
    // ILU(0) level sets: level(I) = 1 + max level of the lower neighbours
    std::vector<int> lvl(N, 0);
    int nL = 0;
    for (long I = 0; I < N; ++I) {
        int l = 0;
        for (int k = T.rowPtr[I]; k < T.diagOff[I]; ++k) l = std::max(l, lvl[T.colIdx[k]] + 1);
        lvl[I] = l;
        nL = std::max(nL, l + 1);
    }
    T.levelPtr.assign(nL + 1, 0);
    for (long I = 0; I < N; ++I) T.levelPtr[lvl[I] + 1]++;
    for (int l = 0; l < nL; ++l) T.levelPtr[l + 1] += T.levelPtr[l];
    T.levelRows.resize(N);
    std::vector<int> fill(T.levelPtr.begin(), T.levelPtr.end() - 1);
    for (long I = 0; I < N; ++I) T.levelRows[fill[lvl[I]]++] = (int32_t)I;
}

}  // namespace omodels
