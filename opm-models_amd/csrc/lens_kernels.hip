// lens_kernels.hip -- element-centred assembly of the two-phase immiscible
// lens problem on gfx950: one thread per element computes the element's
// local residual (fluxes over its faces in intersection order, storage,
// boundary segments, volume-specific scaling) with derivatives w.r.t. its own
// primaries and scatters them by column, as FvBaseLinearizer::linearizeElement_
// does (fvbaselinearizer.hh:597-626): r[I], A[I][I] and A[J][I] for each
// neighbour J.  Every block has one writer, so no clear pass and no atomics.
#include "lens_physics.h"
#include "omodels_internal.h"
#include "lens_ops.h"

namespace omodels {

namespace {

constexpr int kLensThreads = 256;

__device__ __forceinline__ double cellX(const LensParams& P, int i) { return (i + 0.5) * P.dx; }
__device__ __forceinline__ double cellY(const LensParams& P, int j) { return (j + 0.5) * P.dy; }

// ImmiscibleLocalResidual::computeStorage: phi S rho (immisciblelocalresidual.hh:75-103)
__global__ __launch_bounds__(kLensThreads) void k_lens_storage(LensParams P, const double* __restrict__ pv,
                                                               const uint8_t* __restrict__ inLens,
                                                               double* __restrict__ st)
{
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= (long)P.nx * P.ny) return;
    LensIQ q;
    lensIQ(P, inLens[I] != 0, pv[2 * I], pv[2 * I + 1], q);
    for (int a = 0; a < 2; ++a) st[2 * I + a] = ((q.S[a] * P.porosity) * q.rho[a]).v;
}

__global__ __launch_bounds__(kLensThreads) void k_lens_assemble(LensParams P, const double* __restrict__ pv,
    const uint8_t* __restrict__ inLens, const double* __restrict__ st1, double dt, double* __restrict__ res,
    double* __restrict__ jac, const int* __restrict__ diagOff, const int* __restrict__ nbOff)
{
    const int nx = P.nx, ny = P.ny;
    const long N = (long)nx * ny;
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const int i = (int)(I % nx), j = (int)(I / nx);
    LensIQ in;
    lensIQ(P, inLens[I] != 0, pv[2 * I], pv[2 * I + 1], in);
    const double KI = inLens[I] ? P.kLens : P.kOuter;
    const double posI[2] = {cellX(P, i), cellY(P, j)};
    E2 r0[2] = {e2(0.0), e2(0.0)};
    E2 rn[4][2];
    int face[4];
    int nn = 0;
    // YaspGrid intersection order: -x, +x, -y, +y
    const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const int ni = i + di[f], nj = j + dj[f];
        if (!(ni >= 0 && ni < nx && nj >= 0 && nj < ny)) continue;
        const double nrm[2] = {(double)di[f], (double)dj[f]};
        const double posF[2] = {posI[0] + 0.5 * P.dx * di[f], posI[1] + 0.5 * P.dy * dj[f]};
        const long J = (long)ni + (long)nx * nj;
        LensIQv ex;
        lensIQv(P, inLens[J] != 0, pv[2 * J], pv[2 * J + 1], ex);
        const double KJ = inLens[J] ? P.kLens : P.kOuter;
        const double posJ[2] = {cellX(P, ni), cellY(P, nj)};
        E2 flux[2];
        lensInteriorFlux(P, lensHarmonicMean(KI, KJ), in, ex, posI, posJ, posF, nrm, flux);
        const double area = 1.0 * (di[f] != 0 ? P.dy : P.dx);
        for (int e = 0; e < 2; ++e) {
            flux[e] = flux[e] * area;
            r0[e] = r0[e] + flux[e];
            rn[nn][e] = e2(0.0) - flux[e];
        }
        face[nn++] = f;
    }
    // storage with the cached storage of the previous time level
    const double fac = P.V / dt;
    for (int e = 0; e < 2; ++e) {
        const E2 st = (in.S[e] * P.porosity) * in.rho[e];
        E2 t = st - st1[2 * I + e];
        t = t * fac;
        r0[e] = r0[e] + t;
        r0[e] = r0[e] - 0.0 * P.V;   // source term: zero for this problem
    }
    // boundary segments
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const int ni = i + di[f], nj = j + dj[f];
        if (ni >= 0 && ni < nx && nj >= 0 && nj < ny) continue;
        const double nrm[2] = {(double)di[f], (double)dj[f]};
        const double posF[2] = {posI[0] + 0.5 * P.dx * di[f], posI[1] + 0.5 * P.dy * dj[f]};
        E2 values[2];
        lensBoundaryRate(P, KI, in, posI, posF, nrm, values);
        const double area = di[f] != 0 ? P.dy : P.dx;
        for (int e = 0; e < 2; ++e) r0[e] = r0[e] + values[e] * (area * 1.0);
    }
    // volume-specific residual, then the column scatter
    for (int e = 0; e < 2; ++e) r0[e] = r0[e] / P.V;
    res[2 * I] = r0[0].v;
    res[2 * I + 1] = r0[1].v;
    {
        double* b = jac + 4L * diagOff[I];
        b[0] = r0[0].d0; b[1] = r0[0].d1; b[2] = r0[1].d0; b[3] = r0[1].d1;
    }
    for (int k = 0; k < nn; ++k) {
        const E2 a0 = rn[k][0] / P.V, a1 = rn[k][1] / P.V;
        double* b = jac + 4L * nbOff[(long)face[k] * N + I];
        b[0] = a0.d0; b[1] = a0.d1; b[2] = a1.d0; b[3] = a1.d1;
    }
}

// NewtonMethod::update_ (newtonmethod.hh:723-731): x <- x - dx; counts non-finite updates
__global__ __launch_bounds__(kLensThreads) void k_lens_update(long m, double* __restrict__ pv,
                                                              const double* __restrict__ dx, int* __restrict__ bad)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const double d = dx[k];
    const bool fin = isfinite(d);
    if (fin) pv[k] = pv[k] - d;
    const unsigned long long ballot = __ballot(!fin);
    if (ballot && (threadIdx.x & 63) == 0) atomicAdd(bad, __popcll(ballot));
}

}  // namespace

void launchLensStorage(const LensParams& P, const double* pv, const uint8_t* inLens, double* st, hipStream_t s)
{
    const long N = (long)P.nx * P.ny;
    hipLaunchKernelGGL(k_lens_storage, dim3((unsigned)((N + kLensThreads - 1) / kLensThreads)), dim3(kLensThreads),
                       0, s, P, pv, inLens, st);
}

void launchLensAssemble(const LensParams& P, const double* pv, const uint8_t* inLens, const double* st1, double dt,
                        double* res, double* jac, const int* diagOff, const int* nbOff, hipStream_t s)
{
    const long N = (long)P.nx * P.ny;
    hipLaunchKernelGGL(k_lens_assemble, dim3((unsigned)((N + kLensThreads - 1) / kLensThreads)), dim3(kLensThreads),
                       0, s, P, pv, inLens, st1, dt, res, jac, diagOff, nbOff);
}

void launchLensUpdate(long m, double* pv, const double* dx, int* bad, hipStream_t s)
{
    hipLaunchKernelGGL(k_lens_update, dim3((unsigned)((m + kLensThreads - 1) / kLensThreads)), dim3(kLensThreads), 0,
                       s, m, pv, dx, bad);
}

}  // namespace omodels
