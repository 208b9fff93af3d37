// bo_kernels.hip -- black-oil assembly, storage cache and Newton-update
// kernels for gfx950.
//
// Assembly (TpfaLinearizer::linearize_, tpfalinearizer.hh:711-866) runs one
// thread per cell.  A cell owns r[I], the diagonal block A[I][I] and the
// column blocks A[J][I] of its neighbours (the reference's single-writer rule,
// tpfalinearizer.hh:765-770), so every Jacobian block is written exactly once
// per linearization and no clearing pass or atomics are needed.
//
// The kernel is tiled over (x, y) columns of the grid and marches in z: the
// value-only intensive quantities of a 2D tile plus its one-cell halo are
// staged in LDS for the current plane, and the planes below and above come
// from registers of the same thread, so each cell's neighbour quantities are
// evaluated about (1 + halo/tile) times instead of six times.
#include "omodels_internal.h"

namespace omodels {

// ---------------------------------------------------------------------------
// Assembly.  Block = TX x TY threads, one (i, j) column each, k = 0..nz-1.
constexpr int kTX = 32;
constexpr int kTY = 8;

struct IqLds {
    double p, rho[3], lam[3], invB[3], Rs;
};

__device__ __forceinline__ void toLds(const IqV& q, IqLds& l)
{
    l.p = q.p;
    for (int a = 0; a < 3; ++a) { l.rho[a] = q.rho[a]; l.lam[a] = q.lam[a]; l.invB[a] = q.invB[a]; }
    l.Rs = q.Rs;
}
__device__ __forceinline__ void fromLds(const IqLds& l, IqV& q)
{
    q.p = l.p;
    for (int a = 0; a < 3; ++a) { q.rho[a] = l.rho[a]; q.lam[a] = l.lam[a]; q.invB[a] = l.invB[a]; }
    q.Rs = l.Rs;
}

__device__ __forceinline__ void loadIqV(const BoFluid& f, const double* __restrict__ pv,
                                        const uint8_t* __restrict__ mean, long c, IqV& q)
{
    iqValues(f, pv[3 * c], pv[3 * c + 1], pv[3 * c + 2], mean[c], q);
}

// The neighbour J's side of the diffusive and dispersive terms of one face:
// So * 1/Bo of J (the value half of the own-cell saturation formula), the
// face's diffusivity and dispersivity (stored at the transmissibility's index)
// and the mean of the two cells' oil velocities of the previous linearization.
template <bool Diff, bool Disp>
__device__ __forceinline__ DiffFace diffFace(const BoDiff& dif, const BoDisp& dsp, const double* __restrict__ pv,
                                             const uint8_t* __restrict__ mean, long I, long J, const IqV& q,
                                             long face)
{
    DiffFace df{0.0, 0.0, 0.0, 0.0};
    if constexpr (Diff) {
        const double sg = (mean[J] == kMeanSg) ? pv[3 * J + 2] : 0.0;
        const double so = (1.0 - pv[3 * J]) - sg;
        df.sbEx = so * q.invB[kO];
        df.diffusivity = dif.diff[face];
    }
    if constexpr (Disp) {
        df.dispersivity = dsp.disp[face];
        df.nvAvg = 0.5 * (dsp.nvIn[I] + dsp.nvIn[J]);
    }
    return df;
}

// the cell's oil normVelocityCell: the largest |darcy[oil] / faceArea| over
// its faces (tpfalinearizer.hh:759-761, blackoildispersionmodule.hh:374-380)
__device__ __forceinline__ void maxVelocity(double& nv, const double darcy[3], double area)
{
    const double v = fabs(darcy[kCO] / area);
    nv = nv < v ? v : nv;
}

// Accumulate one face of cell I into its residual / diagonal registers and
// store the column block A[J][I] = 0 - dF/dx_I.
template <bool Diff, bool Disp>
__device__ __forceinline__ void applyFace(const BoFluid& f, const IqE& qi, const IqV& qj, double trans,
                                          double area, double dZg, bool upInTie, double r[3],
                                          double D[9], double* __restrict__ Jout, const BoDiff& dif,
                                          DiffFace df, const BoDisp& dsp, double& nv, double thp,
                                          const FaceTM* tm, const FaceDM* dm, bool acc = false)
{
    E3 fl[3];
    double darcy[3];
    faceFluxT<Disp, Diff, Disp>(f, qi, qj, trans, area, dZg, upInTie, fl, darcy, &dif, df, &dsp, thp, tm, dm);
    if constexpr (Disp) maxVelocity(nv, darcy, area);
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        r[e] += fl[e].v;
        D[3 * e + 0] += fl[e].d0;
        D[3 * e + 1] += fl[e].d1;
        D[3 * e + 2] += fl[e].d2;
    }
    double blk[9];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        blk[3 * e + 0] = 0.0 + (-fl[e].d0);
        blk[3 * e + 1] = 0.0 + (-fl[e].d1);
        blk[3 * e + 2] = 0.0 + (-fl[e].d2);
    }
    // a sub-domain's linearization adds the column block into the
    // neighbour's row, which it does not clear (tpfalinearizer.hh:770)
    if (acc) {
#pragma unroll
        for (int t = 0; t < 9; ++t) Jout[t] += blk[t];
    } else {
#pragma unroll
        for (int t = 0; t < 9; ++t) Jout[t] = blk[t];
    }
}

// One (x, y) tile of columns over the planes of one z chunk.  Shared by the
// single-reservoir kernel and the ensemble kernel (per-instance pointers).
// Diff / Disp: with the diffusive / dispersive face term.  Geo: depths,
// volumes and threshold pressures from the caller (geo) instead of the
// structured grid's own; the volume ties of the upwind decision are then
// decided by the volumes, as the reference's Vin / Vex do.
template <bool Diff, bool Disp, bool Geo>
__device__ __forceinline__ void assembleTile(const BoFluid& f, const BoGrid& G,
    const double* __restrict__ pv, const uint8_t* __restrict__ mean,
    const double* __restrict__ poro, const double* __restrict__ trans,
    const double* __restrict__ st1, double storefac, double* __restrict__ res,
    double* __restrict__ jac, const int* __restrict__ diagOff, const int* __restrict__ nbOff, int kChunk,
    int bx, int by, int bz, IqLds (*plane)[kTX + 2], const BoDiff& dif, const BoDisp& dsp, const BoGeo& geo)
{
    const int tx = threadIdx.x % kTX, ty = threadIdx.x / kTX;
    const int i = bx * kTX + tx;
    const int j = by * kTY + ty;
    const bool own = (i < G.nx) && (j < G.ny);
    const long nxy = (long)G.nx * G.ny;
    const long N = nxy * G.nz;

    // halo cells of this tile handled by this thread (perimeter of the tile)
    // index h in [0, 2*(kTX+kTY)+4): enumerate the ring of the (kTX+2)x(kTY+2) box
    const int ringLen = 2 * (kTX + 2) + 2 * kTY;

    // this block marches over planes [k0, k1)
    const int k0 = bz * kChunk;
    const int k1 = min(G.nz, k0 + kChunk);
    IqV below, cur, above;     // own column: k-1, k, k+1 (values)
    bool haveBelow = false;
    if (own) {
        loadIqV(f, pv, mean, (long)i + (long)G.nx * j + nxy * k0, cur);
        if (k0 > 0) { loadIqV(f, pv, mean, (long)i + (long)G.nx * j + nxy * (k0 - 1), below); haveBelow = true; }
    }

    for (int k = k0; k < k1; ++k) {
        const long I = (long)i + (long)G.nx * j + nxy * k;
        // stage plane k: interior from registers, ring computed here
        if (own) toLds(cur, plane[ty + 1][tx + 1]);
        for (int h = threadIdx.x; h < ringLen; h += kTX * kTY) {
            int li, lj;
            if (h < kTX + 2) { li = h; lj = 0; }
            else if (h < 2 * (kTX + 2)) { li = h - (kTX + 2); lj = kTY + 1; }
            else if (h < 2 * (kTX + 2) + kTY) { li = 0; lj = 1 + (h - 2 * (kTX + 2)); }
            else { li = kTX + 1; lj = 1 + (h - 2 * (kTX + 2) - kTY); }
            const int gi = bx * kTX + li - 1, gj = by * kTY + lj - 1;
            if (gi >= 0 && gi < G.nx && gj >= 0 && gj < G.ny) {
                IqV q;
                loadIqV(f, pv, mean, (long)gi + (long)G.nx * gj + nxy * k, q);
                toLds(q, plane[lj][li]);
            }
        }
        // the plane above, for the +z face (own column only)
        if (own && k + 1 < G.nz) loadIqV(f, pv, mean, I + nxy, above);
        __syncthreads();

        // a sub-domain's linearization (Geo form with a domain mask) only
        // evaluates the domain's cells
        bool act = own;
        if constexpr (Geo) act = act && (!geo.domain || geo.domain[I]);
        const bool acc = Geo && geo.domain != nullptr;
        if (act) {
            IqE qi;
            const double sw = pv[3 * I], po = pv[3 * I + 1], x2 = pv[3 * I + 2];
            const int mI = mean[I];
            iqEval(f, sw, po, x2, mI, qi);
            double r[3] = {0.0, 0.0, 0.0};
            double nvo = 0.0;
            double D[9];
#pragma unroll
            for (int t = 0; t < 9; ++t) D[t] = 0.0;
            const double zI = Geo ? geo.depth[I] : G.dz * ((double)k + 0.5);
            // the neighbour's depth, the volume tie of the upwind decision and
            // the face's threshold pressure (plain: the grid's own)
            auto zOf = [&](long J, int dk) -> double {
                if constexpr (Geo) return geo.depth[J];
                else return dk == 0 ? zI : G.dz * ((double)(k + dk) + 0.5);
            };
            auto tieOf = [&](long J) -> bool {
                if constexpr (Geo) {
                    const double vi = geo.vol[I], vj = geo.vol[J];
                    return vi > vj || (vi == vj && I < J);
                } else {
                    return I < J;
                }
            };
            auto thpOf = [&](long fi) -> double {
                if constexpr (Geo) return geo.thp[fi];
                else return 0.0;
            };
            // the face's transmissibility multipliers (given geometry only)
            FaceTM tmv{ec(1.0), 1.0};
            if constexpr (Geo) tmv.in = rockTransMult(geo.rock, qi.p);
            auto tmOf = [&](const IqV& q) -> const FaceTM* {
                if constexpr (Geo) {
                    tmv.ex = rockTransMultV(geo.rock, q.p);
                    return &tmv;
                } else {
                    return nullptr;
                }
            };
            // the directional mobility multipliers of the face's axis (given
            // geometry only)
            FaceDM dmv;
            auto dmOf = [&](long J, int ax) -> const FaceDM* {
                if constexpr (Geo) {
                    if (!geo.dirMob) return nullptr;
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        dmv.in[a] = geo.dirMob[(3 * ax + a) * N + I];
                        dmv.ex[a] = geo.dirMob[(3 * ax + a) * N + J];
                    }
                    return &dmv;
                } else {
                    return nullptr;
                }
            };
            // intersection order: -x, +x, -y, +y, -z, +z
            if (i > 0) {
                IqV q; fromLds(plane[ty + 1][tx], q);
                applyFace<Diff, Disp>(f, qi, q, trans[I - 1], G.area[0], (zI - zOf(I - 1, 0)) * G.g, tieOf(I - 1), r, D,
                                jac + 9L * nbOff[0 * N + I], dif, diffFace<Diff, Disp>(dif, dsp, pv, mean, I, I - 1, q, I - 1), dsp,
                                nvo, thpOf(I - 1), tmOf(q), dmOf(I - 1, 0), acc);
            }
            if (i + 1 < G.nx) {
                IqV q; fromLds(plane[ty + 1][tx + 2], q);
                applyFace<Diff, Disp>(f, qi, q, trans[I], G.area[0], (zI - zOf(I + 1, 0)) * G.g, tieOf(I + 1), r, D,
                                jac + 9L * nbOff[1 * N + I], dif, diffFace<Diff, Disp>(dif, dsp, pv, mean, I, I + 1, q, I), dsp,
                                nvo, thpOf(I), tmOf(q), dmOf(I + 1, 0), acc);
            }
            if (j > 0) {
                IqV q; fromLds(plane[ty][tx + 1], q);
                applyFace<Diff, Disp>(f, qi, q, trans[N + I - G.nx], G.area[1], (zI - zOf(I - G.nx, 0)) * G.g,
                                tieOf(I - G.nx), r, D, jac + 9L * nbOff[2 * N + I], dif,
                                diffFace<Diff, Disp>(dif, dsp, pv, mean, I, I - G.nx, q, N + I - G.nx), dsp, nvo,
                                thpOf(N + I - G.nx), tmOf(q), dmOf(I - G.nx, 1), acc);
            }
            if (j + 1 < G.ny) {
                IqV q; fromLds(plane[ty + 2][tx + 1], q);
                applyFace<Diff, Disp>(f, qi, q, trans[N + I], G.area[1], (zI - zOf(I + G.nx, 0)) * G.g, tieOf(I + G.nx),
                                r, D, jac + 9L * nbOff[3 * N + I], dif,
                                diffFace<Diff, Disp>(dif, dsp, pv, mean, I, I + G.nx, q, N + I), dsp, nvo, thpOf(N + I), tmOf(q),
                                dmOf(I + G.nx, 1), acc);
            }
            if (haveBelow) {
                applyFace<Diff, Disp>(f, qi, below, trans[2 * N + I - nxy], G.area[2], (zI - zOf(I - nxy, -1)) * G.g,
                                tieOf(I - nxy), r, D, jac + 9L * nbOff[4 * N + I], dif,
                                diffFace<Diff, Disp>(dif, dsp, pv, mean, I, I - nxy, below, 2 * N + I - nxy), dsp, nvo,
                                thpOf(2 * N + I - nxy), tmOf(below), dmOf(I - nxy, 2), acc);
            }
            if (k + 1 < G.nz) {
                applyFace<Diff, Disp>(f, qi, above, trans[2 * N + I], G.area[2], (zI - zOf(I + nxy, 1)) * G.g,
                                tieOf(I + nxy), r, D, jac + 9L * nbOff[5 * N + I], dif,
                                diffFace<Diff, Disp>(dif, dsp, pv, mean, I, I + nxy, above, 2 * N + I), dsp, nvo,
                                thpOf(2 * N + I), tmOf(above), dmOf(I + nxy, 2), acc);
            }
            // accumulation term (tpfalinearizer.hh:776-825)
            {
                E3 st[3];
                if constexpr (Geo) storageEval(f, qi, rockPorosity(geo.rock, I, poro[I], qi.p), st);
                else storageEval(f, qi, poro[I], st);
                const double sf = Geo ? geo.vol[I] / geo.dt : storefac;
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    double rv = st[e].v;
                    rv -= st1[3 * I + e];
                    rv *= sf;
                    r[e] += rv;
                    D[3 * e + 0] += st[e].d0 * sf;
                    D[3 * e + 1] += st[e].d1 * sf;
                    D[3 * e + 2] += st[e].d2 * sf;
                }
            }
#pragma unroll
            for (int e = 0; e < 3; ++e) res[3 * I + e] = r[e];
            if constexpr (Disp) dsp.nvOut[I] = nvo;
            double* Dout = jac + 9L * diagOff[I];
#pragma unroll
            for (int t = 0; t < 9; ++t) Dout[t] = D[t];
        }
        __syncthreads();
        below = cur;
        haveBelow = true;
        cur = above;
    }
}

// The fluid tables come through a device pointer: passed by value they fill
// the kernel arguments, and the compiler keeps them in scalar registers that
// then spill into vector registers; read through the pointer they are scalar
// loads where used, which leaves the kernel room for two waves per SIMD.
// The bound of two waves holds the variants with the diffusive and dispersive
// terms there too, at the price of a few spilled registers (faster than one
// wave per SIMD without spills).
template <bool Diff, bool Disp, bool Geo>
__global__ __launch_bounds__(kTX * kTY, 2) void k_bo_assemble(const BoFluid* __restrict__ fp, BoGrid G,
    const double* __restrict__ pv, const uint8_t* __restrict__ mean,
    const double* __restrict__ poro, const double* __restrict__ trans,
    const double* __restrict__ st1, double storefac, double* __restrict__ res,
    double* __restrict__ jac, const int* __restrict__ diagOff, const int* __restrict__ nbOff, int kChunk,
    BoDiff dif, BoDisp dsp, BoGeo geo)
{
    __shared__ IqLds plane[kTY + 2][kTX + 2];
    assembleTile<Diff, Disp, Geo>(*fp, G, pv, mean, poro, trans, st1, storefac, res, jac, diagOff, nbOff, kChunk,
                                  blockIdx.x, blockIdx.y, blockIdx.z, plane, dif, dsp, geo);
}

// The source term of one cell: problem.source(globI) is one rate per cell,
// the sum of the cell's sources in the list's order, then times -V
// (tpfalinearizer.hh:827-835).  grp = {first, count} into order.
__device__ __forceinline__ void cellSourceTerm(const BoFluid& f, const IqE& qi, int2 grp, const int* __restrict__ order,
                                               const int* __restrict__ type, const double* __restrict__ a,
                                               const double* __restrict__ b, double V, E3 t[3])
{
    const int s0 = order[grp.x];
    E3 q[3];
    sourceEval(f, qi, type[s0], a[s0], b[s0], V, q);
    for (int k = 1; k < grp.y; ++k) {
        const int s = order[grp.x + k];
        E3 q2[3];
        sourceEval(f, qi, type[s], a[s], b[s], V, q2);
#pragma unroll
        for (int e = 0; e < 3; ++e) q[e] = q[e] + q2[e];
    }
#pragma unroll
    for (int e = 0; e < 3; ++e) t[e] = q[e] * (-V);
}

// Cell-wise well sources (tpfalinearizer.hh:827-841), added after the flux
// and accumulation terms exactly as the reference's per-cell order does; one
// thread per cell with sources.
__global__ void k_bo_sources(BoFluid f, BoGrid G, const double* __restrict__ pv,
                             const uint8_t* __restrict__ mean, int ngroups, const int2* __restrict__ groups,
                             const int* __restrict__ order, const int* __restrict__ cell,
                             const int* __restrict__ type, const double* __restrict__ a,
                             const double* __restrict__ b, double* __restrict__ res,
                             double* __restrict__ jac, const int* __restrict__ diagOff,
                             const double* __restrict__ vol, const uint8_t* __restrict__ domain)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    const int2 grp = groups[g];
    const long I = cell[order[grp.x]];
    if (domain && !domain[I]) return;      // a sub-domain: its cells' sources only
    const double V = vol ? vol[I] : G.V;
    IqE qi;
    iqEval(f, pv[3 * I], pv[3 * I + 1], pv[3 * I + 2], mean[I], qi);
    E3 t[3];
    cellSourceTerm(f, qi, grp, order, type, a, b, V, t);
    double* D = jac + 9L * diagOff[I];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        res[3 * I + e] += t[e].v;
        D[3 * e + 0] += t[e].d0;
        D[3 * e + 1] += t[e].d1;
        D[3 * e + 2] += t[e].d2;
    }
}

// The same source terms as blocks, not added: block s holds the residual
// contribution (3 values, already times -V) and its derivatives w.r.t. the
// cell's primaries (3x3, row-major) of source s.  With separate sparse source
// terms these are what the built-in well model adds after the cell loop
// (tpfalinearizer.hh:842-846) in place of the cell-wise pass (:827-841).
__global__ void k_bo_source_blocks(BoFluid f, BoGrid G, const double* __restrict__ pv,
                                   const uint8_t* __restrict__ mean, int nsrc, const int* __restrict__ cell,
                                   const int* __restrict__ type, const double* __restrict__ a,
                                   const double* __restrict__ b, double* __restrict__ outRes,
                                   double* __restrict__ outDiag, const double* __restrict__ vol)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsrc) return;
    const long I = cell[s];
    const double V = vol ? vol[I] : G.V;
    IqE qi;
    iqEval(f, pv[3 * I], pv[3 * I + 1], pv[3 * I + 2], mean[I], qi);
    E3 q[3];
    sourceEval(f, qi, type[s], a[s], b[s], V, q);
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const E3 t = q[e] * (-V);
        outRes[3 * s + e] = t.v;
        outDiag[9 * s + 3 * e + 0] = t.d0;
        outDiag[9 * s + 3 * e + 1] = t.d1;
        outDiag[9 * s + 3 * e + 2] = t.d2;
    }
}

// Sparse source terms (wellModel().addReservoirSourceTerms, tpfalinearizer.hh:
// 843-846): residual[cell] += r and A[cell][cell] += D for every entry, in the
// given order.  One wave: lane c < 12 owns component c of every entry, so
// entries for the same cell are added one after another.
__global__ void k_add_sparse_sources(int n, const int* __restrict__ cell, const double* __restrict__ r,
                                     const double* __restrict__ d, double* __restrict__ res,
                                     double* __restrict__ jac, const int* __restrict__ diagOff)
{
    const int c = threadIdx.x;
    if (c >= 12) return;
    for (int k = 0; k < n; ++k) {
        const long I = cell[k];
        if (c < 3) res[3 * I + c] += r[3L * k + c];
        else jac[9L * diagOff[I] + (c - 3)] += d[9L * k + (c - 3)];
    }
}

// The auxiliary module's additions (linearizeAuxiliaryEquations,
// tpfalinearizer.hh:254-281): one wave; lane q < 9 owns entry q of every
// block and lane 9 + e entry e of every residual row, each walking the list in
// order, so entries that meet in one block add in the given order.
__global__ void k_add_auxiliary(int nb, const int* __restrict__ boff, const double* __restrict__ blocks, int nr,
                                const int* __restrict__ rrow, const double* __restrict__ rvals,
                                double* __restrict__ jac, double* __restrict__ res)
{
    const int c = threadIdx.x;
    if (c < 9) {
        for (int k = 0; k < nb; ++k) jac[9L * boff[k] + c] += blocks[9L * k + c];
    } else if (c < 12) {
        const int e = c - 9;
        for (int k = 0; k < nr; ++k) res[3L * rrow[k] + e] += rvals[3L * k + e];
    }
}

// The auxiliary DOFs' Newton update (newtonmethod.hh:704-708):
// next = current - update; a non-finite update is counted in counters[1]
// (the reference rejects any non-finite update, :677-678).
__global__ void k_aux_update(long n, double* __restrict__ x, const double* __restrict__ d,
                             int* __restrict__ counters)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double u = d[i];
    if (!isfinite(u)) atomicAdd(&counters[1], 1);
    x[i] = x[i] - u;
}

// Face rates of the current solution (TpfaLinearizer::updateFlowsInfo,
// tpfalinearizer.hh:647-705): for every cell I and every neighbour J in
// intersection order, the values of the face's component fluxes as the
// residual takes them (FLOWS) and the phases' reservoir-volume rates (FLORES),
// [entry][3] with entries in the pattern's neighbour order.  One thread per
// cell, the assembly's face evaluation.
__global__ void k_bo_flows(const BoFluid* __restrict__ fp, BoGrid G, const double* __restrict__ pv,
                           const uint8_t* __restrict__ mean,
                           const double* __restrict__ trans, const int* __restrict__ rowPtr,
                           double* __restrict__ flows, double* __restrict__ flores, BoDiff dif, BoDisp dsp,
                           BoGeo geo)
{
    const BoFluid& f = *fp;
    const long nxy = (long)G.nx * G.ny;
    const long N = nxy * G.nz;
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const int i = (int)(I % G.nx), j = (int)((I / G.nx) % G.ny), k = (int)(I / nxy);
    IqE qi;
    iqEval(f, pv[3 * I], pv[3 * I + 1], pv[3 * I + 2], mean[I], qi);
    const double zI = geo.depth ? geo.depth[I] : G.dz * ((double)k + 0.5);
    long e = rowPtr[I] - I;     // first neighbour entry of I
    auto face = [&](long J, long fi, double area, double zJ, bool upInTie) {
        IqV qj;
        iqValues(f, pv[3 * J], pv[3 * J + 1], pv[3 * J + 2], mean[J], qj);
        E3 fl[3];
        double darcy[3];
        double thp = 0.0;
        FaceTM tmv{ec(1.0), 1.0};
        const FaceTM* tm = nullptr;
        FaceDM dmv;
        const FaceDM* dm = nullptr;
        if (geo.depth) {
            zJ = geo.depth[J];
            const double vi = geo.vol[I], vj = geo.vol[J];
            upInTie = vi > vj || (vi == vj && I < J);
            thp = geo.thp[fi];
            tmv.in = rockTransMult(geo.rock, qi.p);
            tmv.ex = rockTransMultV(geo.rock, qj.p);
            tm = &tmv;
            if (geo.dirMob) {
                const long ax = fi / N;     // the face's index: axis * N + lower cell
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    dmv.in[a] = geo.dirMob[(3 * ax + a) * N + I];
                    dmv.ex[a] = geo.dirMob[(3 * ax + a) * N + J];
                }
                dm = &dmv;
            }
        }
        const double dZg = (zI - zJ) * G.g;
        if (dif.diff && dsp.disp)
            faceFluxT<true, true, true>(f, qi, qj, trans[fi], area, dZg, upInTie, fl, darcy, &dif,
                                        diffFace<true, true>(dif, dsp, pv, mean, I, J, qj, fi), &dsp, thp, tm, dm);
        else if (dif.diff)
            faceFluxT<true, true>(f, qi, qj, trans[fi], area, dZg, upInTie, fl, darcy, &dif,
                                  diffFace<true, false>(dif, dsp, pv, mean, I, J, qj, fi), nullptr, thp, tm, dm);
        else if (dsp.disp)
            faceFluxT<true, false, true>(f, qi, qj, trans[fi], area, dZg, upInTie, fl, darcy, &dif,
                                         diffFace<false, true>(dif, dsp, pv, mean, I, J, qj, fi), &dsp, thp, tm,
                                         dm);
        else
            faceFluxT<true>(f, qi, qj, trans[fi], area, dZg, upInTie, fl, darcy, nullptr,
                            DiffFace{0.0, 0.0, 0.0, 0.0}, nullptr, thp, tm, dm);
        for (int c = 0; c < 3; ++c) {
            if (flows) flows[3 * e + c] = fl[c].v;
            if (flores) flores[3 * e + c] = darcy[c];
        }
        ++e;
    };
    if (i > 0) face(I - 1, I - 1, G.area[0], zI, false);
    if (i + 1 < G.nx) face(I + 1, I, G.area[0], zI, true);
    if (j > 0) face(I - G.nx, N + I - G.nx, G.area[1], zI, false);
    if (j + 1 < G.ny) face(I + G.nx, N + I, G.area[1], zI, true);
    if (k > 0) face(I - nxy, 2 * N + I - nxy, G.area[2], G.dz * ((double)(k - 1) + 0.5), false);
    if (k + 1 < G.nz) face(I + nxy, 2 * N + I, G.area[2], G.dz * ((double)(k + 1) + 0.5), true);
}

// Output quantities of the current solution, one field after the other
// ([kOutFields][N]): the VtkMultiPhaseModule fields pressure_, saturation_,
// density_ and mobility_<phase> and porosity (vtkmultiphasemodule.hh:306-323),
// then VtkBlackOilModule's R_s, B_w, B_o, B_g and the gas meaning
// (vtkblackoilmodule.hh:276-301).  Phases in the order water, oil, gas.
__global__ void k_bo_output(const BoFluid* __restrict__ fp, long N, const double* __restrict__ pv,
                            const uint8_t* __restrict__ mean, const double* __restrict__ poro,
                            double* __restrict__ out, BoRock rock)
{
    const BoFluid& f = *fp;
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    IqE q;
    iqEval(f, pv[3 * I], pv[3 * I + 1], pv[3 * I + 2], mean[I], q);
    const int ph[3] = {kW, kO, kG};
    for (int a = 0; a < 3; ++a) {
        out[(0 + a) * N + I] = q.p.v;                  // capillary pressure is zero
        out[(3 + a) * N + I] = q.S[ph[a]].v;
        out[(6 + a) * N + I] = q.rho[ph[a]].v;
        out[(9 + a) * N + I] = q.lam[ph[a]].v;
    }
    out[12 * N + I] = rockPorosityV(rock, I, poro[I], pv[3 * I + 1]);
    out[13 * N + I] = q.Rs.v;
    for (int a = 0; a < 3; ++a) out[(14 + a) * N + I] = 1.0 / q.invB[ph[a]].v;
    out[17 * N + I] = (double)mean[I];
}

void launchBoOutput(const BoFluid* fdev, long N, const double* pv, const uint8_t* mean, const double* poro,
                    double* out, hipStream_t s, const BoRock* rock)
{
    if (N <= 0) return;
    hipLaunchKernelGGL(k_bo_output, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, fdev, N, pv, mean, poro,
                       out, rock ? *rock : BoRock{});
}

// Storage of the current solution: the cache of time level 1 at the first
// Newton iteration (recycleFirstIterationStorage, tpfalinearizer.hh:786-812).
__global__ void k_bo_storage(BoFluid f, long N, const double* __restrict__ pv,
                             const uint8_t* __restrict__ mean, const double* __restrict__ poro,
                             double* __restrict__ st, BoRock rock)
{
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    double s[3];
    storageValues(f, pv[3 * I], pv[3 * I + 1], pv[3 * I + 2], mean[I], rockPorosityV(rock, I, poro[I], pv[3 * I + 1]),
                  s);
    st[3 * I] = s[0]; st[3 * I + 1] = s[1]; st[3 * I + 2] = s[2];
}

// ---------------------------------------------------------------------------
// Newton update: BlackOilNewtonMethod::updatePrimaryVariables_
// (blackoilnewtonmethod.hh:226-427) and adaptPrimaryVariables
// (blackoilprimaryvariables.hh:549-909) for the Sw/Po/{Sg,Rs} meanings.
__global__ void k_bo_update(BoFluid f, long N, double* __restrict__ pv, uint8_t* __restrict__ mean,
                            uint8_t* __restrict__ wasSw, const double* __restrict__ delta,
                            NewtonUpdateParams prm, int* __restrict__ counters)
{
    const long I = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    const double d0 = delta[3 * I], d1 = delta[3 * I + 1], d2 = delta[3 * I + 2];
    if (!isfinite(d0) || !isfinite(d1) || !isfinite(d2)) atomicOr(&counters[1], 1);
    const double cur[3] = {pv[3 * I], pv[3 * I + 1], pv[3 * I + 2]};
    const int cm = mean[I];
    double dSw = 0.0, dSo = 0.0, dSg = 0.0;
    dSw = d0;
    dSo -= dSw;
    if (cm == kMeanSg) { dSg = d2; dSo -= dSg; }
    double maxSat = fmax(fabs(dSg), fabs(dSo));
    maxSat = fmax(maxSat, fabs(dSw));
    double satAlpha = 1.0;
    if (maxSat > prm.dsMax) satAlpha = prm.dsMax / maxSat;
    double nx[3];
    {
        double dl = d0 * satAlpha;
        nx[0] = cur[0] - dl;
    }
    {
        double dl = d1;
        if (fabs(dl) > prm.dpMaxRel * cur[1]) {
            const double sgn = (double)((dl > 0.0) - (dl < 0.0));
            dl = sgn * prm.dpMaxRel * cur[1];
        }
        double v = cur[1] - dl;
        v = v < -1e99 ? -1e99 : (v > 1e99 ? 1e99 : v);
        nx[1] = v;
    }
    {
        double dl = d2;
        if (cm == kMeanSg) dl *= satAlpha;
        else if (dl > cur[2]) dl = cur[2];
        nx[2] = cur[2] - dl;
    }
    // adaptPrimaryVariables
    const double eps = wasSw[I] ? prm.oscThresh : 0.0;
    int m = cm;
    bool changed = false;
    const double sw = nx[0];
    const double sg = (m == kMeanSg) ? nx[2] : 0.0;
    if (sw >= 1.0) {
        nx[0] = fmin(1.0, sw);
        nx[2] = 0.0;
        changed = (m != kMeanSg);
        m = kMeanSg;
    } else if (m == kMeanSg) {
        const double s = 1.0 - sw;
        if (sg < -eps && s > 0.0) {
            m = kMeanRs;
            const double rsSat = tabEval(f.rsSat, nx[1]);
            nx[2] = fmin(1.79769313486231570815e+308 / 2, rsSat);
            changed = true;
        }
    } else {
        const double rsSat = tabEval(f.rsSat, nx[1]);
        if (nx[2] > fmin(1.79769313486231570815e+308 / 2, rsSat * (1.0 + eps))) {
            m = kMeanSg;
            nx[2] = 0.0;
            changed = true;
        }
    }
    pv[3 * I] = nx[0]; pv[3 * I + 1] = nx[1]; pv[3 * I + 2] = nx[2];
    mean[I] = (uint8_t)m;
    wasSw[I] = changed ? 1 : 0;
    // per-wave vote, one atomic per wavefront
    const unsigned long long ball = __ballot(changed);
    if ((threadIdx.x % warpSize) == 0 && ball) atomicAdd(&counters[0], __popcll(ball));
}

// ---------------------------------------------------------------------------
// Host launchers.
void launchBoAssemble(const BoFluid* fdev, const BoGrid& G, const double* pv, const uint8_t* mean,
                      const double* poro, const double* trans, const double* st1, double storefac,
                      double* res, double* jac, const int* diagOff, const int* nbOff, hipStream_t s,
                      const BoDiff* dif, const BoDisp* dsp, const BoGeo* geo)
{
    const int gx = (G.nx + kTX - 1) / kTX, gy = (G.ny + kTY - 1) / kTY;
    // enough workgroups to fill the chip: split the z march into chunks
    int gz = (2048 + gx * gy - 1) / (gx * gy);
    gz = gz < 1 ? 1 : (gz > G.nz ? G.nz : gz);
    const int kChunk = (G.nz + gz - 1) / gz;
    gz = (G.nz + kChunk - 1) / kChunk;
    const BoDiff d = (dif && dif->diff) ? *dif : BoDiff{};
    const BoDisp e = (dsp && dsp->disp) ? *dsp : BoDisp{};
    const BoGeo g = (geo && geo->depth) ? *geo : BoGeo{};
    const dim3 grid(gx, gy, gz), block(kTX * kTY);
#define OM_ASSEMBLE(DIFF, DISP, GEO)                                                                          \
    hipLaunchKernelGGL((k_bo_assemble<DIFF, DISP, GEO>), grid, block, 0, s, fdev, G, pv, mean, poro, trans, st1, \
                       storefac, res, jac, diagOff, nbOff, kChunk, d, e, g)
    if (g.depth) {
        if (d.diff && e.disp) OM_ASSEMBLE(true, true, true);
        else if (d.diff) OM_ASSEMBLE(true, false, true);
        else if (e.disp) OM_ASSEMBLE(false, true, true);
        else OM_ASSEMBLE(false, false, true);
    } else {
        if (d.diff && e.disp) OM_ASSEMBLE(true, true, false);
        else if (d.diff) OM_ASSEMBLE(true, false, false);
        else if (e.disp) OM_ASSEMBLE(false, true, false);
        else OM_ASSEMBLE(false, false, false);
    }
#undef OM_ASSEMBLE
}

void launchBoSources(const BoFluid& f, const BoGrid& G, const double* pv, const uint8_t* mean,
                     int ngroups, const int2* groups, const int* order, const int* cell, const int* type,
                     const double* a, const double* b, double* res, double* jac, const int* diagOff, hipStream_t s,
                     const double* vol, const uint8_t* domain)
{
    if (ngroups <= 0) return;
    hipLaunchKernelGGL(k_bo_sources, dim3((ngroups + 63) / 64), dim3(64), 0, s, f, G, pv, mean, ngroups, groups,
                       order, cell, type, a, b, res, jac, diagOff, vol, domain);
}

void launchBoSourceBlocks(const BoFluid& f, const BoGrid& G, const double* pv, const uint8_t* mean,
                          int nsrc, const int* cell, const int* type, const double* a, const double* b,
                          double* outRes, double* outDiag, hipStream_t s, const double* vol)
{
    if (nsrc <= 0) return;
    hipLaunchKernelGGL(k_bo_source_blocks, dim3((nsrc + 63) / 64), dim3(64), 0, s, f, G, pv, mean, nsrc,
                       cell, type, a, b, outRes, outDiag, vol);
}

// Boundary terms after the cell loop and the sources (tpfalinearizer.hh:
// 849-865): one thread per cell with conditions, its faces in face order,
// each face's term added to the residual and the diagonal block.
__global__ void k_bo_boundary(const BoFluid* __restrict__ fp, BoGrid G, const double* __restrict__ pv,
                              const uint8_t* __restrict__ mean, BoBc bc, const double* __restrict__ depth,
                              double* __restrict__ res, double* __restrict__ jac, const int* __restrict__ diagOff,
                              double* __restrict__ flows)
{
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t >= bc.nGroups) return;
    const BoFluid& f = *fp;
    const int2 grp = bc.groups[t];
    const long I = bc.cell[grp.x];
    const long nxy = (long)G.nx * G.ny;
    const double zI = depth ? depth[I] : G.dz * ((double)(I / nxy) + 0.5);
    IqE qi;
    iqEval(f, pv[3 * I], pv[3 * I + 1], pv[3 * I + 2], mean[I], qi);
    double r[3] = {0.0, 0.0, 0.0}, D[9] = {};
    double* Dout = nullptr;
    if (!flows) {        // flows: only the faces' terms are written (res, jac, diagOff unused)
        Dout = jac + 9L * diagOff[I];
#pragma unroll
        for (int e = 0; e < 3; ++e) r[e] = res[3 * I + e];
#pragma unroll
        for (int q = 0; q < 9; ++q) D[q] = Dout[q];
    }
    for (int q = 0; q < grp.y; ++q) {
        const int b = grp.x + q;
        const int type = bc.type[b];
        if (type == 0) continue;
        const int ax = bc.face[b] / 2;
        const double area = G.area[ax];
        E3 out[3];
        if (type == 1) {
#pragma unroll
            for (int e = 0; e < 3; ++e) out[e] = ec(bc.rate[3 * b + e] * area);
        } else {
            IqV qb;
            iqValues(f, bc.state[3 * b], bc.state[3 * b + 1], bc.state[3 * b + 2], bc.meaning[b], qb);
            boundaryFreeFlux(f, qi, qb, bc.trans[b], area, (zI - bc.zFace[b]) * G.g, out);
        }
        if (flows) {     // the face's term as FLOWS records it (tpfalinearizer.hh:700-703)
#pragma unroll
            for (int e = 0; e < 3; ++e) flows[3L * b + e] = out[e].v;
            continue;
        }
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            r[e] += out[e].v;
            D[3 * e + 0] += out[e].d0;
            D[3 * e + 1] += out[e].d1;
            D[3 * e + 2] += out[e].d2;
        }
    }
    if (flows) return;
#pragma unroll
    for (int e = 0; e < 3; ++e) res[3 * I + e] = r[e];
#pragma unroll
    for (int q = 0; q < 9; ++q) Dout[q] = D[q];
}

void launchBoBoundary(const BoFluid* fp, const BoGrid& G, const double* pv, const uint8_t* mean, const BoBc& bc,
                      const double* depth, double* res, double* jac, const int* diagOff, hipStream_t s,
                      double* flows)
{
    if (bc.nGroups <= 0) return;
    hipLaunchKernelGGL(k_bo_boundary, dim3((unsigned)((bc.nGroups + 127) / 128)), dim3(128), 0, s, fp, G, pv, mean, bc,
                       depth, res, jac, diagOff, flows);
}

void launchBoFlows(const BoFluid* fdev, const BoGrid& G, const double* pv, const uint8_t* mean, const double* trans,
                   const int* rowPtr, double* flows, double* flores, hipStream_t s, const BoDiff* dif,
                   const BoDisp* dsp, const BoGeo* geo)
{
    const long N = (long)G.nx * G.ny * G.nz;
    if (N <= 0) return;
    hipLaunchKernelGGL(k_bo_flows, dim3((unsigned)((N + 127) / 128)), dim3(128), 0, s, fdev, G, pv, mean, trans, rowPtr,
                       flows, flores, dif ? *dif : BoDiff{}, dsp ? *dsp : BoDisp{},
                       (geo && geo->depth) ? *geo : BoGeo{});
}

// resetSystem_(domain) (tpfalinearizer.hh:370-379): the domain's rows of the
// residual and the Jacobian set to zero; one thread per row
__global__ void k_clear_rows(int n, const int* __restrict__ cells, const int* __restrict__ rowPtr,
                             double* __restrict__ jac, double* __restrict__ res)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long I = cells[t];
    for (int e = 0; e < 3; ++e) res[3 * I + e] = 0.0;
    for (long k = 9L * rowPtr[I]; k < 9L * rowPtr[I + 1]; ++k) jac[k] = 0.0;
}

void launchClearRows(int n, const int* cells, const int* rowPtr, double* jac, double* res, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_clear_rows, dim3((n + 255) / 256), dim3(256), 0, s, n, cells, rowPtr, jac, res);
}

void launchAddAuxiliary(int nb, const int* blockOff, const double* blocks, int nr, const int* resRow,
                        const double* resVals, double* jac, double* res, hipStream_t s)
{
    if (nb <= 0 && nr <= 0) return;
    hipLaunchKernelGGL(k_add_auxiliary, dim3(1), dim3(64), 0, s, nb, blockOff, blocks, nr, resRow, resVals, jac,
                       res);
}

void launchAuxUpdate(long n, double* x, const double* delta, int* counters, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_aux_update, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, x, delta, counters);
}

void launchAddSparseSources(int n, const int* cell, const double* r, const double* d, double* res,
                            double* jac, const int* diagOff, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_add_sparse_sources, dim3(1), dim3(64), 0, s, n, cell, r, d, res, jac, diagOff);
}

void launchBoStorage(const BoFluid& f, long N, const double* pv, const uint8_t* mean,
                     const double* poro, double* st, hipStream_t s, const BoRock* rock)
{
    hipLaunchKernelGGL(k_bo_storage, dim3((N + 255) / 256), dim3(256), 0, s, f, N, pv, mean, poro, st,
                       rock ? *rock : BoRock{});
}

void launchBoUpdate(const BoFluid& f, long N, double* pv, uint8_t* mean, uint8_t* wasSw,
                    const double* delta, const NewtonUpdateParams& prm, int* counters, hipStream_t s)
{
    hipLaunchKernelGGL(k_bo_update, dim3((N + 255) / 256), dim3(256), 0, s, f, N, pv, mean, wasSw,
                       delta, prm, counters);
}

}  // namespace omodels

// ===========================================================================
// Ensemble kernels: every block finds its instance through a table, so one
// launch serves all realizations of a batch (per-instance pointers, grids and
// time steps in EnsInst).
#include "ensemble.h"

namespace omodels {

// the fluid through a device pointer and two waves per SIMD, as in k_bo_assemble
__global__ __launch_bounds__(kTX * kTY, 2) void k_e_assemble(const BoFluid* __restrict__ fp, const EnsInst* __restrict__ insts,
    const int2* __restrict__ table, const int* __restrict__ active)
{
    __shared__ IqLds plane[kTY + 2][kTX + 2];
    const int2 t = table[blockIdx.x];
    if (!active[t.x]) return;
    const EnsInst& I = insts[t.x];
    const int bx = t.y % I.gx, by = (t.y / I.gx) % I.gy, bz = t.y / (I.gx * I.gy);
    assembleTile<false, false, false>(*fp, I.G, I.pv0, I.mean0, I.poro, I.trans, I.st1, I.storefac, I.res, I.jac,
                                      I.diagOff, I.nbOff, I.kChunk, bx, by, bz, plane, BoDiff{}, BoDisp{}, BoGeo{});
}

// the cell-wise source pass of every instance: table entries (instance,
// group), one thread per cell with sources
__global__ void k_e_sources(BoFluid f, const EnsInst* __restrict__ insts, const int2* __restrict__ table, int n,
                            const int* __restrict__ active)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int2 t = table[q];
    if (!active[t.x]) return;
    const EnsInst& I = insts[t.x];
    const int2 grp = I.srcGroups[t.y];
    const long c = I.srcCell[I.srcOrder[grp.x]];
    IqE qi;
    iqEval(f, I.pv0[3 * c], I.pv0[3 * c + 1], I.pv0[3 * c + 2], I.mean0[c], qi);
    E3 v[3];
    cellSourceTerm(f, qi, grp, I.srcOrder, I.srcType, I.srcA, I.srcB, I.G.V, v);
    double* D = I.jac + 9L * I.diagOff[c];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        I.res[3 * c + e] += v[e].v;
        D[3 * e + 0] += v[e].d0;
        D[3 * e + 1] += v[e].d1;
        D[3 * e + 2] += v[e].d2;
    }
}

__global__ __launch_bounds__(kEnsChunk) void k_e_storage(BoFluid f, const EnsInst* __restrict__ insts,
    const int2* __restrict__ table, const int* __restrict__ refresh)
{
    const int2 t = table[blockIdx.x];
    if (!refresh[t.x]) return;
    const EnsInst& I = insts[t.x];
    const long c = (long)t.y * kEnsChunk + threadIdx.x;
    if (c >= I.N) return;
    double s[3];
    storageValues(f, I.pv0[3 * c], I.pv0[3 * c + 1], I.pv0[3 * c + 2], I.mean0[c], I.poro[c], s);
    I.st1[3 * c] = s[0]; I.st1[3 * c + 1] = s[1]; I.st1[3 * c + 2] = s[2];
}

// Newton update of the instances that solved (same cell rule as k_bo_update)
__global__ __launch_bounds__(kEnsChunk) void k_e_update(BoFluid f, const EnsInst* __restrict__ insts,
    const int2* __restrict__ table, const int* __restrict__ active, NewtonUpdateParams prm)
{
    const int2 t = table[blockIdx.x];
    if (!active[t.x]) return;
    const EnsInst& I = insts[t.x];
    const long c = (long)t.y * kEnsChunk + threadIdx.x;
    const bool valid = c < I.N;
    bool changed = false;
    if (valid) {
        const double* delta = I.dv[kVx];
        double* pv = I.pv0;
        const double d0 = delta[3 * c], d1 = delta[3 * c + 1], d2 = delta[3 * c + 2];
        if (!isfinite(d0) || !isfinite(d1) || !isfinite(d2)) atomicOr(&I.counters[1], 1);
        const double cur[3] = {pv[3 * c], pv[3 * c + 1], pv[3 * c + 2]};
        const int cm = I.mean0[c];
        double dSw = 0.0, dSo = 0.0, dSg = 0.0;
        dSw = d0;
        dSo -= dSw;
        if (cm == kMeanSg) { dSg = d2; dSo -= dSg; }
        double maxSat = fmax(fabs(dSg), fabs(dSo));
        maxSat = fmax(maxSat, fabs(dSw));
        double satAlpha = 1.0;
        if (maxSat > prm.dsMax) satAlpha = prm.dsMax / maxSat;
        double nx[3];
        nx[0] = cur[0] - d0 * satAlpha;
        {
            double dl = d1;
            if (fabs(dl) > prm.dpMaxRel * cur[1]) {
                const double sgn = (double)((dl > 0.0) - (dl < 0.0));
                dl = sgn * prm.dpMaxRel * cur[1];
            }
            double v = cur[1] - dl;
            v = v < -1e99 ? -1e99 : (v > 1e99 ? 1e99 : v);
            nx[1] = v;
        }
        {
            double dl = d2;
            if (cm == kMeanSg) dl *= satAlpha;
            else if (dl > cur[2]) dl = cur[2];
            nx[2] = cur[2] - dl;
        }
        const double eps = I.wasSw[c] ? prm.oscThresh : 0.0;
        int m = cm;
        const double sw = nx[0];
        const double sg = (m == kMeanSg) ? nx[2] : 0.0;
        if (sw >= 1.0) {
            nx[0] = fmin(1.0, sw);
            nx[2] = 0.0;
            changed = (m != kMeanSg);
            m = kMeanSg;
        } else if (m == kMeanSg) {
            const double s = 1.0 - sw;
            if (sg < -eps && s > 0.0) {
                m = kMeanRs;
                nx[2] = fmin(1.79769313486231570815e+308 / 2, tabEval(f.rsSat, nx[1]));
                changed = true;
            }
        } else {
            const double rsSat = tabEval(f.rsSat, nx[1]);
            if (nx[2] > fmin(1.79769313486231570815e+308 / 2, rsSat * (1.0 + eps))) {
                m = kMeanSg;
                nx[2] = 0.0;
                changed = true;
            }
        }
        pv[3 * c] = nx[0]; pv[3 * c + 1] = nx[1]; pv[3 * c + 2] = nx[2];
        I.mean0[c] = (uint8_t)m;
        I.wasSw[c] = changed ? 1 : 0;
    }
    const unsigned long long ball = __ballot(changed);
    if ((threadIdx.x % warpSize) == 0 && ball) atomicAdd(&I.counters[0], __popcll(ball));
}

void launchEnsAssemble(const BoFluid& f, const BoFluid* fdev, const EnsInst* insts, const int2* table, int nBlocks,
                       const int* active, const int2* srcTable, int nSrc, hipStream_t s)
{
    if (nBlocks > 0)
        hipLaunchKernelGGL(k_e_assemble, dim3(nBlocks), dim3(kTX * kTY), 0, s, fdev, insts, table, active);
    if (nSrc > 0)
        hipLaunchKernelGGL(k_e_sources, dim3((nSrc + 63) / 64), dim3(64), 0, s, f, insts, srcTable, nSrc, active);
}

void launchEnsStorage(const BoFluid& f, const EnsInst* insts, const int2* table, int nChunks, const int* refresh,
                      hipStream_t s)
{
    if (nChunks > 0) hipLaunchKernelGGL(k_e_storage, dim3(nChunks), dim3(kEnsChunk), 0, s, f, insts, table, refresh);
}

void launchEnsUpdate(const BoFluid& f, const EnsInst* insts, const int2* table, int nChunks, const int* active,
                     const NewtonUpdateParams& prm, hipStream_t s)
{
    if (nChunks > 0) hipLaunchKernelGGL(k_e_update, dim3(nChunks), dim3(kEnsChunk), 0, s, f, insts, table, active, prm);
}

void assemblyGrid(const BoGrid& G, int& gx, int& gy, int& gz, int& kChunk, int targetBlocks)
{
    gx = (G.nx + kTX - 1) / kTX;
    gy = (G.ny + kTY - 1) / kTY;
    gz = (targetBlocks + gx * gy - 1) / (gx * gy);
    gz = gz < 1 ? 1 : (gz > G.nz ? G.nz : gz);
    kChunk = (G.nz + gz - 1) / gz;
    gz = (G.nz + kChunk - 1) / kChunk;
}

}  // namespace omodels
