/*
 * oracle.h -- CPU restatement of the opm-models fully-implicit TPFA hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is what the parity tests compare the
 * HIP product library with.  It is linked or loaded only by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, and only
 * as that reference or as the timed CPU baseline -- never as a product path.
 *
 * Every function states the reference file:line it restates (paths relative
 * to the opm-models source tree).  The reference itself cannot be compiled
 * here (dune-*, opm-common, opm-simulators are absent), so this restatement
 * is pinned by closed forms, invariants, finite differences and brute force
 * in tests/test_oracle_*.py.
 */
#ifndef OMODELS_ORACLE_H
#define OMODELS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Structured Cartesian black-oil problem.  Cell index I = i + nx*(j + ny*k),
 * x fastest (YaspGrid lexicographic order).  z points up; gravity is the z
 * component of the gravity vector (-9.81 for the usual downward gravity,
 * opm/models/common/multiphasebaseproblem.hh:398-403).
 * Same memory layout as omodels_blackoil_problem in include/omodels.h. */
typedef struct ob_problem {
    int32_t nx, ny, nz, n_src;
    double dx, dy, dz;
    double gravity;
    const double* perm;      /* [3][N] kx, ky, kz in m^2 (SoA) */
    const double* poro;      /* [N] reference porosity */
    int32_t n_oil, n_gas;
    const double* oil_tab;   /* [4][n_oil]: p, Rs_sat, 1/Bo_sat, mu_o_sat */
    const double* gas_tab;   /* [3][n_gas]: p, 1/Bg, mu_g */
    double oil_comp;         /* undersaturated oil compressibility [1/Pa] */
    double water_pref;       /* water reference pressure [Pa] (Bw(pref) = 1) */
    double water_comp;       /* water compressibility [1/Pa] */
    double water_mu;         /* water viscosity [Pa s] */
    double rho_ref[3];       /* surface densities: water, oil, gas [kg/m^3] */
    const int32_t* src_cell; /* [n_src] */
    const int32_t* src_type; /* 0: mass-rate water injector, 1: BHP producer */
    const double* src_a;     /* injector: water mass rate [kg/s]; producer: WI [m^3] */
    const double* src_b;     /* producer: bottom-hole pressure [Pa] */
    /* The linearizer's per-neighbour inputs given by the caller
     * (tpfalinearizer.hh:443-452), each NULL for the structured grid's own:
     * face_trans [3][N], face I|I+e_axis at axis*N + I (transmissibility);
     * depth [N] (dofCenterDepth; dZg = (depth_I - depth_J) * gravity, the
     * grid's own is dz * (k + 0.5)); volume [N] (dofTotalVolume, > 0);
     * thpres [3][N] (thresholdPressure, >= 0, at the face's index). */
    const double* face_trans;
    const double* depth;
    const double* volume;
    const double* thpres;
    /* rock compressibility and compaction (blackoilintensivequantities.hh:
     * 447-480): rock_comp [N] (NULL: 0), rock_pref [N] (NULL: 1e5 Pa),
     * rock_tab [3][n_rock]: p_o, pore-volume and transmissibility multipliers
     * (n_rock 0: none) */
    const double* rock_comp;
    const double* rock_pref;
    int32_t n_rock;
    const double* rock_tab;
    /* directional mobilities (blackoilintensivequantities.hh:530-550): the
     * relative permeabilities along axis d are m_d kr with dir_mob [3][3][N]
     * (axis, phase water/oil/gas, cell; NULL: none) */
    const double* dir_mob;
    /* boundary conditions (TpfaLinearizer::boundaryInfo_, tpfalinearizer.hh:
     * 476-497; BlackOilLocalResidualTPFA::computeBoundaryFlux,
     * blackoillocalresidualtpfa.hh:507-620): n_bc faces, each a boundary
     * face bc_face (0..5: -x, +x, -y, +y, -z, +z) of cell bc_cell; bc_type 0
     * none, 1 rate (bc_rate [n][3] mass rate per area in equation order oil,
     * water, gas), 2 free (bc_state [n][3] primary variables with
     * bc_meaning [n] give the boundary's fluid state); bc_trans [n] (NULL:
     * the cell's half transmissibility) and bc_zface [n] (NULL: the grid's
     * face depth) */
    int32_t n_bc;
    const int32_t* bc_cell;
    const int32_t* bc_face;
    const int32_t* bc_type;
    const double* bc_rate;
    const double* bc_state;
    const uint8_t* bc_meaning;
    const double* bc_trans;
    const double* bc_zface;
    /* auxiliary equations (BaseAuxiliaryModule, baseauxiliarymodule.hh:
     * 55-121): n_aux degrees of freedom after the N cells (rows N ..
     * N + n_aux - 1 of the system) and the pattern entries their module adds
     * (addNeighbors, tpfalinearizer.hh:503-507): block (aux_row[k],
     * aux_col[k]) for k < n_aux_links; every auxiliary row needs its
     * diagonal among them */
    int32_t n_aux;
    int32_t n_aux_links;
    const int32_t* aux_row;
    const int32_t* aux_col;
} ob_problem;

/* Newton / linear-solver parameters; defaults are the reference's
 * (newtonmethodparameters.hh:30-60, blackoilnewtonmethodparameters.hh:33-63,
 * linalgparameters.hh:35-78, fvbaseparameters.hh:112). */
typedef struct ob_params {
    double newton_tol;          /* NewtonTolerance */
    int32_t newton_max_iter;    /* NewtonMaxIterations (20) */
    int32_t newton_target_iter; /* NewtonTargetIterations (10) */
    double newton_max_error;    /* NewtonMaxError (1e100) */
    double lin_tol;             /* LinearSolverTolerance (1e-3) */
    double lin_abs_tol;         /* LinearSolverAbsTolerance (<0: newton_tol/100) */
    double lin_max_error;       /* LinearSolverMaxError (1e7) */
    int32_t lin_max_iter;       /* LinearSolverMaxIterations (1000) */
    int32_t precond;            /* 0: block ILU0, 1: block Jacobi */
    double prec_relax;          /* PreconditionerRelaxation (1.0) */
    int32_t prec_order;         /* PreconditionerOrder: Jacobi sweeps (0 in the reference) */
    int32_t max_dt_divisions;   /* MaxTimeStepDivisions (10) */
    double dp_max_rel;          /* DpMaxRel (0.3) */
    double ds_max;              /* DsMax (0.2) */
    double privar_osc_thresh;   /* PriVarOscilationThreshold (1e-5) */
    double max_dt;              /* MaxTimeStepSize */
    double min_dt;              /* MinTimeStepSize (0) */
    int32_t solver_float;       /* 1: linear solve in float (LinearSolverScalar=float) */
    int32_t verbose;
    int32_t ilu_order;          /* 0: natural (the reference); 1: two-colour 4x2x8 tiles */
    /* precond = 2: aggregation AMG (in place of ParallelAmgBackend) */
    int32_t amg_coarsen_target; /* coarsest level: at most this many cells (1..32; 32) */
    int32_t amg_pre_smooth;     /* forward Gauss-Seidel sweeps before the coarse correction (>= 1; 2) */
    int32_t amg_post_smooth;    /* backward sweeps after it (2) */
    double amg_prolong_damp;    /* x += damp * P xc (1.6) */
    int32_t amg_agg[3];         /* aggregate box in cells (3, 3, 3) */
} ob_params;

/* Per-Newton-iteration record, the tuple the reference prints at
 * newtonmethod.hh:778-782 plus the linear iteration count. */
typedef struct ob_iter_log {
    int32_t step, attempt, iter, lin_iters;
    int32_t num_switched, lin_converged;
    double error, dt;
} ob_iter_log;

/* Molecular diffusion in the oil phase (BlackOilDiffusionModule,
 * blackoildiffusionmodule.hh:168-238, face term of
 * blackoillocalresidualtpfa.hh:469-479): the oil component and the dissolved
 * gas diffuse along the face's Rs difference. */
typedef struct ob_diffusion {
    double oil_in_oil;              /* coefficient of the oil component in the oil phase [m^2/s] */
    double gas_in_oil;              /* coefficient of the gas component in the oil phase [m^2/s] */
    double mw_oil, mw_gas;          /* molar masses; both > 0: mole-fraction form, else mass fractions */
    const double* face_diffusivity; /* [3][N], face I|I+e_axis at axis*N + I; NULL: from porosity */
} ob_diffusion;

/* Dispersion in the oil phase (BlackOilDispersionModule,
 * blackoildispersionmodule.hh:176-254, face term of
 * blackoillocalresidualtpfa.hh:481-493), driven by each cell's oil velocity
 * of the previous linearization. */
typedef struct ob_dispersion {
    const double* face_dispersivity; /* [3][N] like face_diffusivity, or NULL */
    const double* cell_dispersion;   /* [N] dispersion coefficients [m], used when the face array is NULL */
} ob_dispersion;

/* ---- A1: pattern, neighbour info, block addresses ---- */
int ob_pattern_size(const ob_problem* p, int64_t* nnzb, int64_t* nfaces2);
/* linearizeDomain(domain) (tpfalinearizer.hh:226-245) on the n distinct
 * cells: residual and jac are read and written (rows outside the domain keep
 * their values but for the domain cells' column blocks added into them) */
int ob_linearize_domain(const ob_problem* p, const double* pv, const uint8_t* meaning, const double* storage1,
                        double dt, int32_t n, const int32_t* cells, double* residual, double* jac);
/* the boundary faces' FLOWS entries (tpfalinearizer.hh:688-704) in cell and
 * face order; returns their number (cell, face, flows [n][3] may be NULL) */
int ob_boundary_flows(const ob_problem* p, const double* pv, const uint8_t* meaning, int32_t* cell, int32_t* face,
                      double* flows);
/* linearizeAuxiliaryEquations (tpfalinearizer.hh:254-281): what the
 * auxiliary module's linearize(matrix, residual) adds, in order: block k
 * (row brow[k], column bcol[k], row-major 3x3) into J, entry k of the
 * residual (row rrow[k], 3 values) into r; rc -1 when a block is not in the
 * pattern or a row out of range. */
int ob_add_auxiliary(const ob_problem* p, int32_t nb, const int32_t* brow, const int32_t* bcol,
                     const double* blocks, int32_t nr, const int32_t* rrow, const double* rvals, double* jac,
                     double* residual);
/* ob_simulate with auxiliary DOFs: aux [n_aux][3] their solution(0) (in and
 * out); after every linearization fn(user, pv, meaning, aux, jac, residual,
 * newton_iteration) adds the module's terms (a nonzero return fails the
 * attempt) */
typedef int (*ob_aux_fn)(void* user, const double* pv, const uint8_t* meaning, const double* aux, double* jac,
                         double* residual, int iteration);
int ob_simulate_aux(const ob_problem* p, double* pv, uint8_t* meaning, double* aux, ob_aux_fn fn, void* user,
                    int n_steps, double dt0, const ob_params* prm, double* pv_hist, double* dt_hist,
                    ob_iter_log* log, int log_cap, int32_t* log_len, int nthreads);
int ob_build_pattern(const ob_problem* p, int32_t* row_ptr, int32_t* col_idx,
                     int32_t* diag_off, int32_t* nb_ptr, int32_t* nb_cell,
                     int32_t* nb_off, double* nb_trans, double* nb_area,
                     double* nb_dzg, int32_t* nb_dir);

/* ---- A2-A7: linearizeDomain ---- */
int ob_storage(const ob_problem* p, const double* pv, const uint8_t* meaning,
               double* storage /* [N][3] */);
int ob_linearize(const ob_problem* p, const double* pv, const uint8_t* meaning,
                 const double* storage1, double dt, double* residual /* [N][3] */,
                 double* jac /* [NNZB][3][3] */, int nthreads);
/* face rates (FLOWS, FLORES) per neighbour entry of the pattern, [entries][3] */
int ob_flows(const ob_problem* p, const double* pv, const uint8_t* meaning, double* flows, double* flores);
/* the same with the diffusive and the dispersive face terms (dif, dsp may be
 * NULL).  With dsp: nv_in [N] the cells' oil velocities of the previous
 * linearization, nv_out [N] receives this one's. */
/* reps linearizations of one problem on nthreads threads, the wall time of each */
int ob_linearize_timed(const ob_problem* p, const double* pv, const uint8_t* meaning, const double* storage1,
                       double dt, int nthreads, int reps, double* seconds);
int ob_linearize_ext(const ob_problem* p, const ob_diffusion* dif, const ob_dispersion* dsp, const double* nv_in,
                     double* nv_out, const double* pv, const uint8_t* meaning, const double* storage1, double dt,
                     double* residual, double* jac, int nthreads);
int ob_flows_ext(const ob_problem* p, const ob_diffusion* dif, const ob_dispersion* dsp, const double* nv,
                 const double* pv, const uint8_t* meaning, double* flows, double* flores);
/* source terms as sparse blocks: cells [n_src], residual [n_src][3] (times -V),
 * derivatives [n_src][3][3]; returns n_src */
int ob_source_blocks(const ob_problem* p, const double* pv, const uint8_t* meaning, int32_t* cells,
                     double* residual, double* diag);
/* Intensive quantities of one cell (values and derivatives); used by tests
 * to pin the PVT and the AD propagation. out: [13][4] = (p, rho_w, rho_o,
 * rho_g, lam_w, lam_o, lam_g, invB_w, invB_o, invB_g, Rs, phi, So) x (v,d0,d1,d2) */
int ob_intensive(const ob_problem* p, const double* pv3, int meaning, int cell,
                 double* out);

/* ---- A10-A14: BCRS linear algebra ---- */
int ob_ilu0_factor(int n, int bs, const int32_t* row_ptr, const int32_t* col_idx,
                   double* vals);
int ob_ilu0_apply(int n, int bs, const int32_t* row_ptr, const int32_t* col_idx,
                  const double* lu, const double* d, double* v, double relax);
/* block ILU(0) in the two-colour tile ordering (ilu_order = 1) of an
 * nx x ny x nz 7-point pattern: factor, then v = M^-1 d (times relax) */
int ob_colored_ilu(int nx, int ny, int nz, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                   const double* d, double* v, double relax);
/* ob_bicgstab on the pattern of an nx x ny x nz grid (needed for ilu_order 1) */
int ob_bicgstab_grid(int nx, int ny, int nz, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                     const double* b, double* x, const ob_params* prm, int32_t* iterations);
/* aggregation AMG (precond = 2, also through ob_bicgstab_grid): level
 * dimensions ([levels][3], returns the level count), a level's matrix in the
 * structured layout [7][N_l][9] (slots -z, -y, -x, I, +x, +y, +z), one
 * V-cycle v = M^-1 d; one block Gauss-Seidel sweep of a structured-layout
 * matrix in red-black order (forward / backward, zero: x = 0 before); the
 * in-place dense inverse of the coarsest level */
int ob_amg_dims(int nx, int ny, int nz, const ob_params* prm, int32_t* dims, int cap);
int ob_amg_level_matrix(int nx, int ny, int nz, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                        const ob_params* prm, int level, double* A7);
int ob_amg_apply(int nx, int ny, int nz, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                 const ob_params* prm, const double* d, double* v);
int ob_gs_sweep(int nx, int ny, int nz, const double* A7, const double* b, double* x, int forward, int zero);
int ob_dense_inverse(int n, double* M);
/* block Jacobi of ob_bicgstab (precond = 1): `order` sweeps on v in place */
int ob_jacobi_apply(int n, int bs, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                    const double* d, double* v, int order, double relax);
/* scalar product of the restated BiCGStab: the exact sum of the products
 * x[i]*y[i], rounded once to the nearest double */
double ob_dot(int64_t m, const double* x, const double* y);
int ob_spmv(int n, int bs, const int32_t* row_ptr, const int32_t* col_idx,
            const double* vals, const double* x, double* y);
int ob_bicgstab(int n, int bs, const int32_t* row_ptr, const int32_t* col_idx,
                const double* vals, const double* b, double* x,
                const ob_params* prm, int32_t* iterations);

/* ---- A15-A16: Newton update of one iterate ---- */
int ob_update_primary(const ob_problem* p, const double* cur, const uint8_t* cur_meaning,
                      const double* delta, double* next, uint8_t* next_meaning,
                      uint8_t* was_switched, const ob_params* prm, int32_t* num_switched);

/* ---- A15-A18: full run (time loop + Newton + backend) ---- */
int ob_simulate(const ob_problem* p, double* pv, uint8_t* meaning, int n_steps,
                double dt0, const ob_params* prm, double* pv_hist /* [n_steps][N][3] or NULL */,
                double* dt_hist /* [n_steps] */, ob_iter_log* log, int log_cap,
                int32_t* log_len, double* timers /* [4] linearize, solve, update, total */,
                int nthreads);

int ob_simulate_ext(const ob_problem* p, const ob_diffusion* dif, const ob_dispersion* dsp, double* pv,
                    uint8_t* meaning, int n_steps, double dt0, const ob_params* prm, double* pv_hist,
                    double* dt_hist, ob_iter_log* log, int log_cap, int32_t* log_len, double* timers, int nthreads);

void ob_default_params(ob_params* prm);

/* Two-phase immiscible lens problem on a 2-D structured grid (config 1,
 * tests/problems/lensproblem.hh).  Cell index I = i + nx*j, x fastest; y up.
 * Primary variables per cell: p_w [Pa], S_w.  Same layout as
 * omodels_lens_problem in include/omodels.h. */
typedef struct ol_problem {
    int32_t nx, ny;
    double lx, ly;                                   /* domain size [m] */
    double lens_x0, lens_y0, lens_x1, lens_y1;       /* lens box [m] */
    double k_lens, k_outer;                          /* intrinsic permeability [m^2] */
    double porosity;
    double swr_lens, snr_lens, alpha_lens, n_lens;   /* regularized Van Genuchten, lens */
    double swr_outer, snr_outer, alpha_outer, n_outer;
    double rho_w, rho_n, mu_w, mu_n;                 /* incompressible liquids */
    double gravity;                                  /* y component [m/s^2] */
    double inflow_rate;                              /* DNAPL inflow on the top strip [kg/(m^2 s)] */
    double p_atm;                                    /* reference pressure of the free-flow sides [Pa] */
    double eps;                                      /* geometric tolerance [m] */
} ol_problem;

void ol_default_problem(ol_problem* p);
int ol_num_cells(const ol_problem* p, int64_t* n, int64_t* nnzb);
int ol_pattern(const ol_problem* p, int32_t* row_ptr, int32_t* col_idx, int32_t* diag);
int ol_initial(const ol_problem* p, double* pv);
int ol_storage(const ol_problem* p, const double* pv, double* storage);
int ol_linearize(const ol_problem* p, const double* pv, const double* storage1, double dt,
                 double* residual, double* jac);
int ol_material(const ol_problem* p, int lens, double sw, double* out);
int ol_simulate(const ol_problem* p, double* pv, double t_end, double dt0, const ob_params* prm,
                int max_steps, double* t_hist, double* dt_hist, int32_t* n_steps,
                ob_iter_log* log, int log_cap, int32_t* log_len);

#ifdef __cplusplus
}
#endif
#endif
