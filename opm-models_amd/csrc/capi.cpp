// capi.cpp -- extern "C" boundary of libomodels_amd.so (declared in
// include/omodels.h).  No C++ exception crosses this boundary: every entry
// point maps NumericalProblem to 2, other errors to a negative code, and keeps
// the message in a thread-local string.
#include <algorithm>
#include <cstring>
#include <limits>
#include <memory>
#include <string>

#include "ensemble_host.h"
#include "lens_host.h"
#include "omodels_host.h"
#include "omodels_io.h"

using namespace omodels;

namespace {
thread_local std::string g_lastError;

template <class F>
int guarded(F&& f)
{
    try {
        return f();
    } catch (const NumericalProblem& e) {
        g_lastError = e.what();
        return 2;
    } catch (const std::invalid_argument& e) {
        g_lastError = e.what();
        return -2;
    } catch (const std::exception& e) {
        g_lastError = e.what();
        return -1;
    } catch (...) {
        g_lastError = "unknown error";
        return -1;
    }
}
}  // namespace

struct omodels_model {
    std::unique_ptr<Model> m;
};
struct omodels_server {
    std::unique_ptr<Server> s;
};
struct omodels_lens {
    std::unique_ptr<LensModel> m;
};
struct omodels_solver {
    omodels_model* owner;
    std::unique_ptr<BiCGStabBackend> s;
};

extern "C" {

const char* omodels_last_error(void) { return g_lastError.c_str(); }

void omodels_default_params(omodels_params* p)
{
    std::memset(p, 0, sizeof(*p));
    p->newton_tol = 1e-6;           // ReservoirProblem's NewtonTolerance (reservoirproblem.hh:401)
    p->newton_max_iter = 20;
    p->newton_target_iter = 10;
    p->newton_max_error = 1e100;
    p->lin_tol = 1e-3;
    p->lin_abs_tol = -1.0;
    p->lin_max_error = 1e7;
    p->lin_max_iter = 1000;
    p->precond = 0;
    p->prec_relax = 1.0;
    p->prec_order = 0;
    p->max_dt_divisions = 10;
    p->dp_max_rel = 0.3;
    p->ds_max = 0.2;
    p->privar_osc_thresh = 1e-5;
    p->max_dt = std::numeric_limits<double>::infinity();
    p->min_dt = 0.0;
    p->precision = 0;
    p->verbose = 0;
    p->amg_coarsen_target = 32;
    p->amg_pre_smooth = 2;
    p->amg_post_smooth = 2;
    p->amg_prolong_damp = 1.6;
    p->amg_agg[0] = p->amg_agg[1] = p->amg_agg[2] = 3;
}

int omodels_blackoil_default_fluid(double* oil_tab, int32_t* n_oil, double* gas_tab, int32_t* n_gas,
                                   double* consts)
{
    return guarded([&] {
        std::vector<double> oil, gas;
        int no, ng;
        defaultFluid(oil, no, gas, ng, consts);
        std::copy(oil.begin(), oil.end(), oil_tab);
        std::copy(gas.begin(), gas.end(), gas_tab);
        *n_oil = no; *n_gas = ng;
        return 0;
    });
}

int omodels_blackoil_generate(const omodels_blackoil_spec* spec, double* perm, double* poro, double* pv,
                              uint8_t* meaning, int32_t src_cap, int32_t* n_src, int32_t* src_cell,
                              int32_t* src_type, double* src_a, double* src_b)
{
    return guarded([&] {
        std::vector<double> k, ph, x, a, b;
        std::vector<uint8_t> mn;
        std::vector<int32_t> c, t;
        generate(*spec, k, ph, x, mn, c, t, a, b);
        if ((long)c.size() > src_cap) throw std::invalid_argument("source capacity too small");
        std::copy(k.begin(), k.end(), perm);
        std::copy(ph.begin(), ph.end(), poro);
        std::copy(x.begin(), x.end(), pv);
        std::copy(mn.begin(), mn.end(), meaning);
        std::copy(c.begin(), c.end(), src_cell);
        std::copy(t.begin(), t.end(), src_type);
        std::copy(a.begin(), a.end(), src_a);
        std::copy(b.begin(), b.end(), src_b);
        *n_src = (int32_t)c.size();
        return 0;
    });
}

omodels_model* omodels_model_create(const omodels_blackoil_problem* prob)
{
    omodels_model* h = nullptr;
    int rc = guarded([&] {
        auto m = std::make_unique<Model>(prob);
        h = new omodels_model{std::move(m)};
        return 0;
    });
    return rc == 0 ? h : nullptr;
}

void omodels_model_destroy(omodels_model* m) { delete m; }

int omodels_model_num_cells(const omodels_model* m, int64_t* n, int64_t* nnzb)
{
    if (!m) return -2;
    *n = m->m->N; *nnzb = m->m->nnzb;
    return 0;
}

int omodels_model_set_solution(omodels_model* h, const double* pv, const uint8_t* meaning, int both_levels)
{
    return guarded([&] {
        Model& m = *h->m;
        m.pv0.upload(pv, 3 * m.N, m.stream);
        m.mean0.upload(meaning, m.N, m.stream);
        if (both_levels) m.advanceTimeLevel();
        m.sync();
        return 0;
    });
}

int omodels_model_get_solution(omodels_model* h, double* pv, uint8_t* meaning)
{
    return guarded([&] {
        Model& m = *h->m;
        m.pv0.download(pv, 3 * m.N, m.stream);
        if (meaning) m.mean0.download(meaning, m.N, m.stream);
        m.sync();
        return 0;
    });
}

int omodels_model_set_storage1(omodels_model* h, const double* st)
{
    return guarded([&] {
        Model& m = *h->m;
        if (st) m.st1.upload(st, 3 * m.N, m.stream);
        else launchBoStorage(m.fluid, m.N, m.pv1.p, m.mean1.p, m.poro.p, m.st1.p, m.stream, &m.geo.rock);
        m.sync();
        return 0;
    });
}

int omodels_model_get_storage1(omodels_model* h, double* st)
{
    return guarded([&] {
        Model& m = *h->m;
        m.st1.download(st, 3 * m.N, m.stream);
        m.sync();
        return 0;
    });
}

int omodels_linearizer_pattern(omodels_model* h, int32_t* row_ptr, int32_t* col_idx, int32_t* diag_off,
                               int32_t* nb_ptr, int32_t* nb_cell, int32_t* nb_off, double* nb_trans,
                               double* nb_dzg)
{
    return guarded([&] {
        const Model& m = *h->m;
        const HostTopology& T = m.topo;
        const long N = T.N;
        const bool aux = m.nAux > 0;   // the pattern with the auxiliary module's rows and entries
        const std::vector<int>& rp = aux ? m.xRowPtr : T.rowPtr;
        const std::vector<int>& ci = aux ? m.xColIdx : T.colIdx;
        const std::vector<int>& dg = aux ? m.xDiagOff : T.diagOff;
        const std::vector<int>& nbo = aux ? m.xNbOff : T.nbOff;
        if (row_ptr) std::copy(rp.begin(), rp.end(), row_ptr);
        if (col_idx) std::copy(ci.begin(), ci.end(), col_idx);
        if (diag_off) std::copy(dg.begin(), dg.end(), diag_off);
        // neighbour info in intersection order (-x,+x,-y,+y,-z,+z)
        long k = 0;
        if (nb_ptr) nb_ptr[0] = 0;
        const long nxy = (long)T.nx * T.ny;
        for (long I = 0; I < N; ++I) {
            const long kk = I / nxy;
            for (int d = 0; d < 6; ++d) {
                const int off = nbo[(long)d * N + I];
                if (off < 0) continue;
                const long J = d == 0 ? I - 1 : d == 1 ? I + 1 : d == 2 ? I - T.nx : d == 3 ? I + T.nx
                             : d == 4 ? I - nxy : I + nxy;
                const int ax = d / 2;
                const long face = (d & 1) ? I : J;
                if (nb_cell) nb_cell[k] = (int32_t)J;
                if (nb_off) nb_off[k] = off;
                if (nb_trans) nb_trans[k] = T.trans[(long)ax * N + face];
                if (nb_dzg) {
                    const long kj = J / nxy;
                    const bool own = m.depthHost.empty();
                    const double zI = own ? m.G.dz * ((double)kk + 0.5) : m.depthHost[I];
                    const double zJ = own ? m.G.dz * ((double)kj + 0.5) : m.depthHost[J];
                    nb_dzg[k] = (zI - zJ) * m.G.g;
                }
                ++k;
            }
            if (nb_ptr) nb_ptr[I + 1] = (int32_t)k;
        }
        return 0;
    });
}

int omodels_linearizer_update_discretization(omodels_model* h, const double* trans, const double* depth,
                                             const double* volume, const double* thpres)
{
    return guarded([&] {
        h->m->updateDiscretization(trans, depth, volume, thpres);
        return 0;
    });
}

int omodels_model_set_rock(omodels_model* h, const double* compressibility, const double* ref_pressure,
                           int32_t n_table, const double* table)
{
    return guarded([&] {
        h->m->setRock(compressibility, ref_pressure, n_table, table);
        return 0;
    });
}

int omodels_model_set_boundary_conditions(omodels_model* h, int32_t n, const int32_t* cell, const int32_t* face,
                                          const int32_t* type, const double* mass_rate, const double* state,
                                          const uint8_t* meaning, const double* trans, const double* face_depth)
{
    return guarded([&] {
        h->m->setBoundary(n, cell, face, type, mass_rate, state, meaning, trans, face_depth);
        return 0;
    });
}

int omodels_model_set_auxiliary(omodels_model* h, int32_t n_dofs, int32_t n_links, const int32_t* row,
                                const int32_t* col)
{
    return guarded([&] {
        h->m->setAuxiliary(n_dofs, n_links, row, col);
        return 0;
    });
}

int omodels_model_set_auxiliary_solution(omodels_model* h, const double* x, int both_levels)
{
    return guarded([&] {
        Model& m = *h->m;
        if (m.nAux == 0) throw std::invalid_argument("no auxiliary DOFs");
        if (!x) throw std::invalid_argument("auxiliary solution must not be NULL");
        m.aux0.upload(x, 3 * (size_t)m.nAux, m.stream);
        if (both_levels) m.aux1.upload(x, 3 * (size_t)m.nAux, m.stream);
        m.sync();
        return 0;
    });
}

int omodels_model_get_auxiliary_solution(omodels_model* h, double* x)
{
    return guarded([&] {
        Model& m = *h->m;
        if (m.nAux == 0) throw std::invalid_argument("no auxiliary DOFs");
        m.aux0.download(x, 3 * (size_t)m.nAux, m.stream);
        m.sync();
        return 0;
    });
}

int omodels_model_set_auxiliary_module(omodels_model* h, omodels_aux_linearize_fn fn, void* user)
{
    return guarded([&] {
        Model& m = *h->m;
        m.auxFn = reinterpret_cast<int (*)(void*, void*, int)>(fn);
        m.auxUser = user;
        m.auxHandle = h;
        return 0;
    });
}

int omodels_linearizer_add_auxiliary(omodels_model* h, int32_t n_blocks, const int32_t* block_row,
                                     const int32_t* block_col, const double* blocks, int32_t n_res,
                                     const int32_t* res_row, const double* res)
{
    return guarded([&] {
        Model& m = *h->m;
        if (n_blocks < 0 || n_res < 0) throw std::invalid_argument("auxiliary: counts must be >= 0");
        if ((n_blocks > 0 && (!block_row || !block_col || !blocks)) || (n_res > 0 && (!res_row || !res)))
            throw std::invalid_argument("auxiliary: arrays must not be NULL");
        std::vector<int> off(n_blocks), rr(n_res);
        for (int k = 0; k < n_blocks; ++k) {
            const long o = m.blockOffset(block_row[k], block_col[k]);
            if (o < 0) throw std::invalid_argument("auxiliary: a block outside the pattern");
            off[k] = (int)o;
        }
        for (int k = 0; k < n_res; ++k) {
            if (res_row[k] < 0 || res_row[k] >= m.rows()) throw std::invalid_argument("auxiliary: row out of range");
            rr[k] = res_row[k];
        }
        if (n_blocks == 0 && n_res == 0) return 0;
        DevBuf<int> doff, drr;
        DevBuf<double> db, dr;
        doff.upload(off.data(), off.size(), m.stream);
        db.upload(blocks, 9 * (size_t)n_blocks, m.stream);
        drr.upload(rr.data(), rr.size(), m.stream);
        dr.upload(res, 3 * (size_t)n_res, m.stream);
        m.addAuxiliary(n_blocks, doff.p, db.p, n_res, drr.p, dr.p);
        m.sync();
        return 0;
    });
}

int omodels_linearizer_boundary_flows(omodels_model* h, int32_t* cell, int32_t* face, double* flows)
{
    return guarded([&] { return h->m->boundaryFlows(flows, cell, face); });
}

int omodels_model_set_directional_mobility(omodels_model* h, const double* multipliers)
{
    return guarded([&] {
        h->m->setDirectionalMobility(multipliers);
        return 0;
    });
}

int omodels_linearizer_linearize_domain(omodels_model* h, double dt, int iteration_index)
{
    return guarded([&] {
        if (!(dt > 0)) throw std::invalid_argument("dt must be positive");
        h->m->linearizeDomain(dt, iteration_index);
        h->m->sync();
        return 0;
    });
}

int omodels_linearizer_linearize_subdomain(omodels_model* h, double dt, int iteration_index, int32_t n,
                                           const int32_t* cells)
{
    return guarded([&] {
        if (!(dt > 0)) throw std::invalid_argument("dt must be positive");
        h->m->linearizeSubdomain(dt, iteration_index, n, cells);
        return 0;
    });
}

int omodels_linearizer_get_jacobian(omodels_model* h, double* values)
{
    return guarded([&] {
        Model& m = *h->m;
        m.jac.download(values, 9 * (size_t)m.nnzb, m.stream);
        m.sync();
        return 0;
    });
}

int omodels_linearizer_get_residual(omodels_model* h, double* r)
{
    return guarded([&] {
        Model& m = *h->m;
        m.res.download(r, 3 * (size_t)m.rows(), m.stream);
        m.sync();
        return 0;
    });
}

int omodels_linearizer_erase_matrix(omodels_model* h)
{
    return guarded([&] { h->m->eraseMatrix(); return 0; });
}

int omodels_linearizer_set_separate_sparse_sources(omodels_model* h, int enable)
{
    return guarded([&] { h->m->separateSparseSources = enable != 0; return 0; });
}

int omodels_linearizer_source_blocks(omodels_model* h, int32_t* cells, double* residual, double* diag)
{
    return guarded([&] {
        Model& m = *h->m;
        if (m.nsrc <= 0) return 0;
        m.computeSourceBlocks();
        if (cells) m.srcCell.download(cells, m.nsrc, m.stream);
        if (residual) m.srcRes.download(residual, 3 * (size_t)m.nsrc, m.stream);
        if (diag) m.srcDiag.download(diag, 9 * (size_t)m.nsrc, m.stream);
        m.sync();
        return m.nsrc;
    });
}

int omodels_linearizer_flows(omodels_model* h, double* flows, double* flores)
{
    return guarded([&] {
        Model& m = *h->m;
        const size_t nf = (size_t)(m.topo.rowPtr[m.N] - m.N);
        if (nf == 0 || (!flows && !flores)) return 0;
        DevBuf<double> df, dr;
        if (flows) df.alloc(3 * nf);
        if (flores) dr.alloc(3 * nf);
        m.computeFlows(df.p, dr.p);
        if (flows) df.download(flows, 3 * nf, m.stream);
        if (flores) dr.download(flores, 3 * nf, m.stream);
        m.sync();
        return 0;
    });
}

int omodels_model_set_diffusion(omodels_model* h, const omodels_diffusion* d)
{
    return guarded([&] {
        h->m->setDiffusion(d);
        return 0;
    });
}

int omodels_model_set_dispersion(omodels_model* h, const omodels_dispersion* d)
{
    return guarded([&] {
        h->m->setDispersion(d);
        return 0;
    });
}

int omodels_model_get_cell_velocity(omodels_model* h, double* velocity)
{
    return guarded([&] {
        Model& m = *h->m;
        if (!m.dispersion.disp) throw std::invalid_argument("dispersion is off");
        OM_HIP(hipMemcpyAsync(velocity, m.dispersion.nvIn, m.N * sizeof(double), hipMemcpyDeviceToHost, m.stream));
        m.sync();
        return 0;
    });
}

// ---- output and restart files ----
namespace {
const char* const kOutNames[kOutFields] = {
    "pressure_water", "pressure_oil", "pressure_gas", "saturation_water", "saturation_oil", "saturation_gas",
    "density_water", "density_oil", "density_gas", "mobility_water", "mobility_oil", "mobility_gas",
    "porosity", "R_s", "B_w", "B_o", "B_g", "primary vars meaning gas"};

void checkGrid(const char* path, int32_t nx, int32_t ny, int32_t nz)
{
    if (!path) throw std::invalid_argument("file path must not be NULL");
    if (nx <= 0 || ny <= 0 || nz <= 0) throw std::invalid_argument("grid dimensions must be positive");
}
}  // namespace

int omodels_output_field_count(void) { return kOutFields; }

const char* omodels_output_field_name(int32_t i) { return (i >= 0 && i < kOutFields) ? kOutNames[i] : nullptr; }

int omodels_output_fields(omodels_model* h, double* out)
{
    return guarded([&] {
        Model& m = *h->m;
        if (!out) throw std::invalid_argument("output array must not be NULL");
        DevBuf<double> d;
        d.alloc((size_t)kOutFields * m.N);
        m.timer.begin(kClsOther, m.stream);
        launchBoOutput(m.fluidDev.p, m.N, m.pv0.p, m.mean0.p, m.poro.p, d.p, m.stream, &m.geo.rock);
        m.timer.end(m.stream);
        OM_HIP(hipGetLastError());
        d.download(out, (size_t)kOutFields * m.N, m.stream);
        m.sync();
        return 0;
    });
}

int omodels_restart_write(const char* path, int32_t nx, int32_t ny, int32_t nz, const double* pv,
                          const uint8_t* meaning, double time, int32_t time_step_idx)
{
    return guarded([&] {
        checkGrid(path, nx, ny, nz);
        if (!pv || !meaning) throw std::invalid_argument("solution arrays must not be NULL");
        writeRestart(path, nx, ny, nz, pv, meaning, time, time_step_idx);
        return 0;
    });
}

int omodels_restart_read(const char* path, int32_t nx, int32_t ny, int32_t nz, double* pv, uint8_t* meaning,
                         double* time, int32_t* time_step_idx)
{
    return guarded([&] {
        checkGrid(path, nx, ny, nz);
        if (!pv || !meaning) throw std::invalid_argument("solution arrays must not be NULL");
        int step = 0;
        readRestart(path, nx, ny, nz, pv, meaning, time, &step);
        if (time_step_idx) *time_step_idx = step;
        return 0;
    });
}

int omodels_vtk_write(const char* path, int32_t nx, int32_t ny, int32_t nz, double dx, double dy, double dz,
                      int32_t nfields, const char* const* names, const double* fields)
{
    return guarded([&] {
        checkGrid(path, nx, ny, nz);
        if (nfields < 0 || (nfields > 0 && (!names || !fields))) throw std::invalid_argument("bad field arrays");
        writeVtu(path, nx, ny, nz, dx, dy, dz, nfields, names, fields);
        return 0;
    });
}

int omodels_model_write_restart(omodels_model* h, const char* path, double time, int32_t time_step_idx)
{
    return guarded([&] {
        Model& m = *h->m;
        checkGrid(path, m.G.nx, m.G.ny, m.G.nz);
        std::vector<double> pv(3 * (size_t)m.N);
        std::vector<uint8_t> mean(m.N);
        m.pv0.download(pv.data(), pv.size(), m.stream);
        m.mean0.download(mean.data(), mean.size(), m.stream);
        m.sync();
        writeRestart(path, m.G.nx, m.G.ny, m.G.nz, pv.data(), mean.data(), time, time_step_idx);
        return 0;
    });
}

int omodels_model_read_restart(omodels_model* h, const char* path, double* time, int32_t* time_step_idx)
{
    return guarded([&] {
        Model& m = *h->m;
        checkGrid(path, m.G.nx, m.G.ny, m.G.nz);
        std::vector<double> pv(3 * (size_t)m.N);
        std::vector<uint8_t> mean(m.N);
        int step = 0;
        readRestart(path, m.G.nx, m.G.ny, m.G.nz, pv.data(), mean.data(), time, &step);
        if (time_step_idx) *time_step_idx = step;
        m.pv0.upload(pv.data(), pv.size(), m.stream);
        m.mean0.upload(mean.data(), mean.size(), m.stream);
        m.advanceTimeLevel();
        m.sync();
        return 0;
    });
}

int omodels_model_write_vtk(omodels_model* h, const char* path)
{
    return guarded([&] {
        Model& m = *h->m;
        checkGrid(path, m.G.nx, m.G.ny, m.G.nz);
        std::vector<double> out((size_t)kOutFields * m.N);
        if (int rc = omodels_output_fields(h, out.data())) return rc;
        writeVtu(path, m.G.nx, m.G.ny, m.G.nz, m.G.dx, m.G.dy, m.G.dz, kOutFields, kOutNames, out.data());
        return 0;
    });
}

int omodels_linearizer_add_sparse_sources(omodels_model* h, int32_t n, const int32_t* cells, const double* residual,
                                          const double* diag)
{
    return guarded([&] {
        Model& m = *h->m;
        if (n < 0) throw std::invalid_argument("negative number of sparse source entries");
        if (n == 0) return 0;
        if (!cells || !residual || !diag) throw std::invalid_argument("sparse source arrays must not be NULL");
        for (int k = 0; k < n; ++k)
            if (cells[k] < 0 || cells[k] >= m.N) throw std::invalid_argument("sparse source cell out of range");
        DevBuf<int> dc;
        DevBuf<double> dr, dd;
        dc.upload(cells, n, m.stream);
        dr.upload(residual, 3 * (size_t)n, m.stream);
        dd.upload(diag, 9 * (size_t)n, m.stream);
        m.addSparseSources(n, dc.p, dr.p, dd.p);
        m.sync();
        return 0;
    });
}

omodels_solver* omodels_solver_create(omodels_model* h, const omodels_params* prm)
{
    omodels_solver* s = nullptr;
    int rc = guarded([&] {
        s = new omodels_solver{h, std::make_unique<BiCGStabBackend>(*h->m, *prm)};
        return 0;
    });
    return rc == 0 ? s : nullptr;
}

void omodels_solver_destroy(omodels_solver* s) { delete s; }

int omodels_solver_prepare(omodels_solver* s, const double* values, const double* residual)
{
    return guarded([&] { s->s->prepare(values, residual); return 0; });
}

int omodels_solver_get_residual(omodels_solver* s, double* residual)
{
    return guarded([&] {
        Model& m = *s->owner->m;
        OM_HIP(hipMemcpyAsync(residual, m.res.p, 3 * m.rows() * sizeof(double), hipMemcpyDeviceToHost, m.stream));
        m.sync();
        return 0;
    });
}

int omodels_solver_solve(omodels_solver* s, double* x)
{
    return guarded([&] {
        const int rc = s->s->solve();
        Model& m = *s->owner->m;
        if (x) {
            OM_HIP(hipMemcpyAsync(x, s->s->solution(), 3 * m.rows() * sizeof(double), hipMemcpyDeviceToHost,
                                  m.stream));
            m.sync();
        }
        return rc;
    });
}

int omodels_solver_iterations(const omodels_solver* s) { return s->s->iterations(); }

int omodels_scalar_product(int64_t m, const double* x, const double* y, double* out)
{
    return guarded([&] {
        if (m < 0 || m >= (int64_t(1) << 28)) throw std::invalid_argument("scalar product: 0 <= m < 2^28");
        if (!out || (m > 0 && (!x || !y))) throw std::invalid_argument("scalar product: missing array");
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("no HIP device available");
        hipStream_t st = nullptr;
        OM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); } } g{st};
        DevBuf<double> dx, dy, dout;
        DevBuf<DotScratch> sc;
        dx.upload(x, (size_t)m, st);
        dy.upload(y, (size_t)m, st);
        dout.alloc(1);
        sc.alloc(1);
        OM_HIP(hipMemsetAsync(sc.p, 0, sizeof(DotScratch), st));
        launchScalarProduct((long)m, dx.p, dy.p, sc.p, dout.p, st);
        OM_HIP(hipGetLastError());
        dout.download(out, 1, st);
        OM_HIP(hipStreamSynchronize(st));
        return 0;
    });
}

int omodels_scalar_product_noalloc(int64_t m, const double* x_dev, const double* y_dev, double* out_dev,
                                   void* stream)
{
    return guarded([&] {
        scalarProductNoAlloc((long)m, x_dev, y_dev, out_dev, static_cast<hipStream_t>(stream));
        return 0;
    });
}

int omodels_solver_get_ilu(omodels_solver* s, double* lu)
{
    return guarded([&] { s->s->downloadIlu(lu); return 0; });
}

int omodels_solver_apply_precond(omodels_solver* s, const double* d, double* v)
{
    return guarded([&] { s->s->applyPrecond(d, v); return 0; });
}

int omodels_solver_amg_levels(omodels_solver* s, int32_t* dims, int cap)
{
    return guarded([&] {
        const AmgHierarchy& a = s->s->amg();
        for (int l = 0; l < a.levels() && l < cap; ++l)
            for (int q = 0; q < 3; ++q) dims[3 * l + q] = a.dims(l)[q];
        return a.levels();
    });
}

int omodels_solver_get_amg_level(omodels_solver* s, int level, double* values)
{
    return guarded([&] {
        s->s->downloadAmgLevel(level, values);
        return 0;
    });
}

int omodels_newton_update(omodels_model* h, const omodels_params* prm, const double* delta, int32_t* num_switched)
{
    return guarded([&] {
        Model& m = *h->m;
        DevBuf<double> d;
        d.upload(delta, 3 * m.rows(), m.stream);
        BlackOilNewtonMethod nm(m, *prm);
        int nsw = 0;
        const int rc = nm.update(d.p, &nsw);
        *num_switched = nsw;
        return rc;
    });
}

int omodels_model_get_switched(omodels_model* h, uint8_t* was_switched)
{
    return guarded([&] {
        Model& m = *h->m;
        m.wasSw.download(was_switched, m.N, m.stream);
        m.sync();
        return 0;
    });
}

int omodels_simulate(omodels_model* h, const omodels_params* prm, int n_steps, double dt0, double* pv_hist,
                     double* dt_hist, omodels_iter_log* log, int log_cap, int32_t* log_len,
                     omodels_run_stats* stats)
{
    return guarded([&] {
        std::vector<omodels_iter_log> lg;
        omodels_run_stats st{};
        const int rc = simulate(*h->m, *prm, n_steps, dt0, pv_hist, dt_hist, &lg, &st);
        if (log) {
            const int n = std::min<int>(log_cap, (int)lg.size());
            std::copy(lg.begin(), lg.begin() + n, log);
        }
        if (log_len) *log_len = (int32_t)lg.size();
        if (stats) *stats = st;
        return rc;
    });
}

omodels_server* omodels_server_create(const omodels_params* prm, double hbm_budget_bytes, int max_batch)
{
    omodels_server* h = nullptr;
    int rc = guarded([&] {
        h = new omodels_server{std::make_unique<Server>(*prm, hbm_budget_bytes, max_batch)};
        return 0;
    });
    return rc == 0 ? h : nullptr;
}

void omodels_server_destroy(omodels_server* s) { delete s; }

int omodels_server_submit(omodels_server* s, const omodels_blackoil_problem* p, const double* pv,
                          const uint8_t* meaning, int n_steps, double dt0, double arrival_s)
{
    return guarded([&] { return s->s->submit(p, pv, meaning, n_steps, dt0, arrival_s); });
}

int omodels_server_run(omodels_server* s)
{
    return guarded([&] { s->s->run(); return 0; });
}

int omodels_server_result(omodels_server* s, int id, double* pv, uint8_t* meaning, double* times, double* dt_hist,
                          omodels_run_stats* stats)
{
    return guarded([&] {
        const Server::Request& r = s->s->request(id);
        if (r.status < 0) throw std::invalid_argument("request has not run");
        if (pv) std::copy(r.result.begin(), r.result.end(), pv);
        if (meaning) std::copy(r.resultMean.begin(), r.resultMean.end(), meaning);
        if (times) { times[0] = r.arrival; times[1] = r.admit; times[2] = r.finish; }
        if (dt_hist) std::copy(r.dtHist.begin(), r.dtHist.end(), dt_hist);
        if (stats) *stats = r.stats;
        return r.status;
    });
}

int omodels_server_log(omodels_server* s, int id, omodels_iter_log* log, int cap, int32_t* len)
{
    return guarded([&] {
        const Server::Request& r = s->s->request(id);
        const int n = std::min<int>(cap, (int)r.log.size());
        if (log) std::copy(r.log.begin(), r.log.begin() + n, log);
        if (len) *len = (int32_t)r.log.size();
        return 0;
    });
}

int omodels_server_summary(omodels_server* s, double* wall_s, int64_t* batch_steps, int32_t* max_running)
{
    return guarded([&] {
        if (wall_s) *wall_s = s->s->wall();
        if (batch_steps) *batch_steps = s->s->batchSteps();
        if (max_running) *max_running = s->s->maxRunning();
        return 0;
    });
}

int omodels_server_inject_ilu_wait_error(omodels_server* s)
{
    return guarded([&] {
        s->s->ensemble().raiseIluError();
        return 0;
    });
}

int omodels_server_set_profiling(omodels_server* s, int enable)
{
    return guarded([&] {
        s->s->ensemble().timer.reset();
        s->s->ensemble().timer.enabled = enable > 0;
        s->s->ensemble().timer.every = std::max(1, enable);
        return 0;
    });
}

int omodels_server_kernel_times(omodels_server* s, double* ms, int64_t* counts, int n)
{
    return guarded([&] {
        KernelTimer& t = s->s->ensemble().timer;
        t.collect();
        for (int i = 0; i < n && i < kNumClasses; ++i) { ms[i] = t.estimatedMs(i); counts[i] = t.launched[i]; }
        return 0;
    });
}

int omodels_model_set_profiling(omodels_model* h, int enable)
{
    return guarded([&] {
        h->m->timer.reset();
        h->m->timer.enabled = enable > 0;
        h->m->timer.every = std::max(1, enable);
        return 0;
    });
}

int omodels_model_kernel_times(omodels_model* h, double* ms, int64_t* counts, int n)
{
    return guarded([&] {
        KernelTimer& t = h->m->timer;
        t.collect();
        for (int i = 0; i < n && i < kNumClasses; ++i) { ms[i] = t.estimatedMs(i); counts[i] = t.launched[i]; }
        return 0;
    });
}

// ---------------------------------------------------------------------------
// two-phase immiscible element path (lens problem)

void omodels_lens_default_problem(omodels_lens_problem* p)
{
    std::memset(p, 0, sizeof(*p));
    p->nx = 48; p->ny = 32;                         // CellsX/Y (lensproblem.hh:301-302)
    p->lx = 6.0; p->ly = 4.0;                       // DomainSizeX/Y (:303-304)
    p->lens_x0 = 1.0; p->lens_y0 = 2.0; p->lens_x1 = 4.0; p->lens_y1 = 3.0;   // (:130-148)
    p->k_lens = 9.05e-12; p->k_outer = 4.6e-10;     // (:260-261)
    p->porosity = 0.4;                              // (:370-373)
    p->swr_lens = 0.18; p->snr_lens = 0.0; p->alpha_lens = 0.00045; p->n_lens = 7.3;      // (:246-258)
    p->swr_outer = 0.05; p->snr_outer = 0.0; p->alpha_outer = 0.0037; p->n_outer = 4.7;
    p->rho_w = 1000.0; p->rho_n = 1460.0;           // SimpleH2O / DNAPL liquid densities
    p->mu_w = 1.0e-3; p->mu_n = 5.7e-4;             // and viscosities
    p->gravity = -9.81;                             // (multiphasebaseproblem.hh:398-403)
    p->inflow_rate = 0.04;                          // (:525-530)
    p->p_atm = 1e5;
    p->eps = 3e-6;                                  // (:232)
}

omodels_lens* omodels_lens_create(const omodels_lens_problem* p)
{
    omodels_lens* h = nullptr;
    int rc = guarded([&] {
        h = new omodels_lens{std::make_unique<LensModel>(p)};
        return 0;
    });
    return rc == 0 ? h : nullptr;
}

void omodels_lens_destroy(omodels_lens* m) { delete m; }

int omodels_lens_num_cells(const omodels_lens* h, int64_t* n, int64_t* nnzb)
{
    if (!h) return -2;
    *n = h->m->N; *nnzb = h->m->nnzb;
    return 0;
}

int omodels_lens_pattern(omodels_lens* h, int32_t* row_ptr, int32_t* col_idx, int32_t* diag)
{
    return guarded([&] {
        const LensModel& m = *h->m;
        std::copy(m.hRowPtr.begin(), m.hRowPtr.end(), row_ptr);
        std::copy(m.hColIdx.begin(), m.hColIdx.end(), col_idx);
        if (diag) std::copy(m.hDiag.begin(), m.hDiag.end(), diag);
        return 0;
    });
}

int omodels_lens_initial(omodels_lens* h, double* pv)
{
    return guarded([&] { h->m->initial(pv); return 0; });
}

int omodels_lens_set_solution(omodels_lens* h, const double* pv, int both_levels)
{
    return guarded([&] {
        LensModel& m = *h->m;
        m.pv0.upload(pv, 2 * m.N, m.stream);
        if (both_levels) m.advanceTimeLevel();
        m.sync();
        return 0;
    });
}

int omodels_lens_get_solution(omodels_lens* h, double* pv)
{
    return guarded([&] {
        LensModel& m = *h->m;
        m.pv0.download(pv, 2 * m.N, m.stream);
        m.sync();
        return 0;
    });
}

int omodels_lens_set_storage1(omodels_lens* h, const double* st)
{
    return guarded([&] {
        LensModel& m = *h->m;
        m.st1.upload(st, 2 * m.N, m.stream);
        m.sync();
        return 0;
    });
}

int omodels_lens_linearize(omodels_lens* h, double dt, int iteration_index)
{
    return guarded([&] {
        h->m->linearizeDomain(dt, iteration_index);
        h->m->sync();
        return 0;
    });
}

int omodels_lens_get_jacobian(omodels_lens* h, double* values)
{
    return guarded([&] {
        LensModel& m = *h->m;
        m.jac.download(values, 4 * (size_t)m.nnzb, m.stream);
        m.sync();
        return 0;
    });
}

int omodels_lens_get_residual(omodels_lens* h, double* residual)
{
    return guarded([&] {
        LensModel& m = *h->m;
        m.res.download(residual, 2 * m.N, m.stream);
        m.sync();
        return 0;
    });
}

int omodels_lens_simulate(omodels_lens* h, const omodels_params* prm, double t_end, double dt0, int max_steps,
                          double* t_hist, double* dt_hist, int32_t* n_steps, omodels_iter_log* log, int log_cap,
                          int32_t* log_len, omodels_run_stats* stats)
{
    return guarded([&] {
        std::vector<omodels_iter_log> lg;
        omodels_run_stats st{};
        int ns = 0;
        const int rc = simulateLens(*h->m, *prm, t_end, dt0, max_steps, t_hist, dt_hist, &ns, &lg, &st);
        if (n_steps) *n_steps = ns;
        if (log) {
            const int n = std::min<int>(log_cap, (int)lg.size());
            std::copy(lg.begin(), lg.begin() + n, log);
        }
        if (log_len) *log_len = (int32_t)lg.size();
        if (stats) *stats = st;
        return rc;
    });
}

}  // extern "C"
