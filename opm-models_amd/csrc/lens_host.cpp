// lens_host.cpp -- host side of the two-phase immiscible element path.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <set>
#include <stdexcept>

#include "lens_host.h"
#include "lens_ops.h"

namespace omodels {

namespace {

double vgPc(const LensMat& p, double swe) { return std::pow(std::pow(swe, -1.0 / p.m) - 1.0, 1.0 / p.n) / p.alpha; }
double vgDpc(const LensMat& p, double swe)
{
    return -1.0 / (p.alpha * p.n * p.m) * std::pow(std::pow(swe, -1.0 / p.m) - 1.0, 1.0 / p.n - 1.0) *
           std::pow(swe, -1.0 / p.m - 1.0);
}

// RegularizedVanGenuchtenParams::finalize: the threshold values and slopes
LensMat makeMat(double swr, double snr, double alpha, double n)
{
    if (!(n > 1.0) || !(alpha > 0.0) || !(swr + snr < 1.0)) throw std::invalid_argument("invalid Van Genuchten parameters");
    LensMat p;
    p.swr = swr; p.snr = snr; p.alpha = alpha; p.n = n; p.m = 1.0 - 1.0 / n;
    p.pcLow = vgPc(p, 0.01);
    p.slopeLow = vgDpc(p, 0.01);
    p.pcHigh = vgPc(p, 0.99);
    p.slopeHigh = vgDpc(p, 0.99);
    p.m1 = (0.0 - p.pcHigh) / (1.0 - 0.99) * 2.0;
    return p;
}

}  // namespace

LensParams makeLensParams(const omodels_lens_problem* p)
{
    if (!p || p->nx < 1 || p->ny < 1 || !(p->lx > 0) || !(p->ly > 0)) throw std::invalid_argument("invalid lens grid");
    LensParams P{};
    P.nx = p->nx; P.ny = p->ny;
    P.lx = p->lx; P.ly = p->ly;
    P.dx = p->lx / p->nx; P.dy = p->ly / p->ny; P.V = P.dx * P.dy;
    P.lensX0 = p->lens_x0; P.lensY0 = p->lens_y0; P.lensX1 = p->lens_x1; P.lensY1 = p->lens_y1;
    P.kLens = p->k_lens; P.kOuter = p->k_outer; P.porosity = p->porosity;
    P.lens = makeMat(p->swr_lens, p->snr_lens, p->alpha_lens, p->n_lens);
    P.outer = makeMat(p->swr_outer, p->snr_outer, p->alpha_outer, p->n_outer);
    P.rhoW = p->rho_w; P.rhoN = p->rho_n; P.muW = p->mu_w; P.muN = p->mu_n;
    P.g = p->gravity; P.inflow = p->inflow_rate; P.pAtm = p->p_atm; P.eps = p->eps;
    return P;
}

LensModel::LensModel(const omodels_lens_problem* p, hipStream_t s)
{
    P = makeLensParams(p);
    initSystem(s);
    const int nx = P.nx, ny = P.ny;
    N = (long)nx * ny;
    bs = 2;
    gx = nx; gy = ny; gz = 1;
    // pattern: the 5-point stencil, sorted columns
    hRowPtr.assign(N + 1, 0);
    for (long I = 0; I < N; ++I) {
        const int i = I % nx, j = I / nx;
        hRowPtr[I + 1] = hRowPtr[I] + 1 + (i > 0) + (i + 1 < nx) + (j > 0) + (j + 1 < ny);
    }
    nnzb = hRowPtr[N];
    hColIdx.resize(nnzb);
    hDiag.resize(N);
    for (long I = 0; I < N; ++I) {
        const int i = I % nx, j = I / nx;
        int k = hRowPtr[I];
        if (j > 0) hColIdx[k++] = (int32_t)(I - nx);
        if (i > 0) hColIdx[k++] = (int32_t)(I - 1);
        hDiag[I] = k;
        hColIdx[k++] = (int32_t)I;
        if (i + 1 < nx) hColIdx[k++] = (int32_t)(I + 1);
        if (j + 1 < ny) hColIdx[k++] = (int32_t)(I + nx);
    }
    // block of row J, column I for each face of I (-x, +x, -y, +y): the
    // column-ownership scatter target of FvBaseLinearizer::linearizeElement_
    std::vector<int32_t> nb(4 * N, -1);
    auto blockOf = [&](long row, long col) {
        for (int k = hRowPtr[row]; k < hRowPtr[row + 1]; ++k)
            if (hColIdx[k] == col) return k;
        throw std::logic_error("block not in pattern");
    };
    const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
    for (long I = 0; I < N; ++I) {
        const int i = I % nx, j = I / nx;
        for (int f = 0; f < 4; ++f) {
            const int ni = i + di[f], nj = j + dj[f];
            if (ni < 0 || ni >= nx || nj < 0 || nj >= ny) continue;
            nb[f * N + I] = blockOf((long)ni + (long)nx * nj, I);
        }
    }
    hInLens.resize(N);
    for (long I = 0; I < N; ++I) {
        const double x = (I % nx + 0.5) * P.dx, y = (I / nx + 0.5) * P.dy;
        hInLens[I] = !(x < P.lensX0 - P.eps || x > P.lensX1 + P.eps || y < P.lensY0 - P.eps ||
                       y > P.lensY1 + P.eps);
    }
    // level schedule of the lower triangle (rows of one level are independent)
    std::vector<int> lev(N, 0);
    int nLev = 0;
    for (long I = 0; I < N; ++I) {
        int l = 0;
        for (int k = hRowPtr[I]; k < hDiag[I]; ++k) l = std::max(l, lev[hColIdx[k]] + 1);
        lev[I] = l;
        nLev = std::max(nLev, l + 1);
    }
    std::vector<int> ptr(nLev + 1, 0), rows(N);
    for (long I = 0; I < N; ++I) ptr[lev[I] + 1]++;
    for (int l = 0; l < nLev; ++l) ptr[l + 1] += ptr[l];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (long I = 0; I < N; ++I) rows[fill[lev[I]]++] = (int)I;
    levels.nLevels = nLev;
    levels.ptrHost = ptr;
    levels.rows.upload(rows.data(), rows.size(), stream);

    rowPtr.upload(hRowPtr.data(), hRowPtr.size(), stream);
    colIdx.upload(hColIdx.data(), hColIdx.size(), stream);
    diagOff.upload(hDiag.data(), hDiag.size(), stream);
    nbOff.upload(nb.data(), nb.size(), stream);
    inLens.upload(hInLens.data(), hInLens.size(), stream);
    pv0.alloc(2 * N); pv1.alloc(2 * N); st1.alloc(2 * N); res.alloc(2 * N);
    jac.alloc(4 * (size_t)nnzb);
    bad.alloc(1);
    OM_HIP(hipMemsetAsync(st1.p, 0, 2 * N * sizeof(double), stream));
    OM_HIP(hipMemsetAsync(jac.p, 0, 4 * (size_t)nnzb * sizeof(double), stream));
    std::vector<double> init(2 * N);
    initial(init.data());
    pv0.upload(init.data(), init.size(), stream);
    OM_HIP(hipMemcpyAsync(pv1.p, pv0.p, 2 * N * sizeof(double), hipMemcpyDeviceToDevice, stream));
    sync();
}

void LensModel::initial(double* pv) const
{
    for (long I = 0; I < N; ++I) {
        const double depth = P.ly - (I / P.nx + 0.5) * P.dy;
        pv[2 * I] = P.pAtm - P.rhoW * P.g * depth;
        pv[2 * I + 1] = 1.0;
    }
}

void LensModel::linearizeDomain(double dt, int iterationIndex)
{
    if (!(dt > 0.0)) throw std::invalid_argument("time step size must be positive");
    if (iterationIndex == 0) launchLensStorage(P, pv0.p, inLens.p, st1.p, stream);
    timer.begin(kClsAssemble, stream);
    launchLensAssemble(P, pv0.p, inLens.p, st1.p, dt, res.p, jac.p, diagOff.p, nbOff.p, stream);
    timer.end(stream);
    OM_HIP(hipGetLastError());
}

void LensModel::advanceTimeLevel()
{
    OM_HIP(hipMemcpyAsync(pv1.p, pv0.p, 2 * N * sizeof(double), hipMemcpyDeviceToDevice, stream));
}

void LensModel::restoreTimeLevel()
{
    OM_HIP(hipMemcpyAsync(pv0.p, pv1.p, 2 * N * sizeof(double), hipMemcpyDeviceToDevice, stream));
}

bool LensModel::update(const double* dx)
{
    OM_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), stream));
    launchLensUpdate(2 * N, pv0.p, dx, bad.p, stream);
    int b = 0;
    OM_HIP(hipMemcpyAsync(&b, bad.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    sync();
    return b == 0;
}

int simulateLens(LensModel& m, const omodels_params& prm, double tEnd, double dt0, int maxSteps, double* tHist,
                 double* dtHist, int* nSteps, std::vector<omodels_iter_log>* log, omodels_run_stats* st)
{
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    if (!(dt0 > 0.0) || !(tEnd > 0.0)) throw std::invalid_argument("t_end and dt0 must be positive");
    BiCGStabBackend solver(m, prm);
    double t = 0.0;
    double dt = std::min(dt0, tEnd);
    int steps = 0;
    auto finished = [&]() {
        const double eps = std::max(std::abs(t), dt) * std::numeric_limits<double>::epsilon() * 1e3;
        return t * (1.0 + eps) >= tEnd;
    };
    m.advanceTimeLevel();
    while (!finished() && steps < maxSteps) {
        bool stepOk = false;
        int numIter = 0;
        for (int attempt = 0; attempt < prm.max_dt_divisions; ++attempt) {
            // NewtonMethod::apply
            numIter = 0;
            double error = 1e100, lastError = 1e100;
            bool failed = false;
            auto converged = [&]() { return error <= prm.newton_tol; };
            auto proceed = [&]() {
                if (numIter < 1) return true;
                if (converged()) return false;
                if (numIter >= prm.newton_max_iter) return error * 4.0 < lastError;
                return true;
            };
            while (proceed()) {
                const auto a = clk::now();
                m.linearizeDomain(dt, numIter);
                lastError = error;
                error = m.maxAbsResidual();
                const auto b = clk::now();
                if (st) { st->t_linearize += std::chrono::duration<double>(b - a).count(); st->linearizations++; }
                if (!(error <= prm.newton_max_error)) { failed = true; break; }
                if (!proceed()) {
                    if (log) log->push_back({steps, attempt, numIter, 0, 0, 1, error, dt});
                    ++numIter;
                    break;
                }
                solver.prepare();
                const int rc = solver.solve();
                const auto c = clk::now();
                if (st) { st->t_solve += std::chrono::duration<double>(c - b).count(); st->linear_iters += solver.iterations(); }
                if (rc != 0) {
                    if (log) log->push_back({steps, attempt, numIter, solver.iterations(), 0, 0, error, dt});
                    failed = true;
                    break;
                }
                const bool fin = m.update(solver.solution());
                if (st) { st->t_update += std::chrono::duration<double>(clk::now() - c).count(); st->newton_iters++; }
                if (log) log->push_back({steps, attempt, numIter, solver.iterations(), 0, 1, error, dt});
                if (prm.verbose)
                    printf("step %d attempt %d iter %d error %.6e lin %d\n", steps, attempt, numIter, error,
                           solver.iterations());
                if (!fin) { failed = true; break; }
                ++numIter;
            }
            if (!failed && converged()) { stepOk = true; break; }
            if (st) st->failed_attempts++;
            numIter = 2 * prm.newton_target_iter;
            m.restoreTimeLevel();
            const double ndt = dt / 2.0;
            if (ndt < prm.min_dt) break;
            dt = ndt;
        }
        if (!stepOk) {
            m.sync();
            if (nSteps) *nSteps = steps;
            return 3;
        }
        m.advanceTimeLevel();
        t += dt;
        if (tHist) tHist[steps] = t;
        if (dtHist) dtHist[steps] = dt;
        ++steps;
        const int tgt = prm.newton_target_iter;
        double sug;
        if (numIter > tgt) sug = dt / (1.0 + double(numIter - tgt) / tgt);
        else sug = dt * (1.0 + double(tgt - numIter) / tgt / 1.2);
        double dtNext = std::min(prm.max_dt, sug);
        const double remain = std::max(0.0, tEnd - t);
        if (dtNext < remain && remain < dtNext * 2) dtNext = remain / 2 * 1.01;
        dt = std::min(remain, dtNext);
    }
    m.sync();
    m.timer.collect();
    if (nSteps) *nSteps = steps;
    if (st) st->t_total = std::chrono::duration<double>(clk::now() - t0).count();
    return 0;
}

}  // namespace omodels
