/*
 * omodels.h -- C ABI of the MI355X-native opm-models hot path.
 *
 * The library keeps the three interfaces the reference's Newton loop sits on:
 *   - Linearizer            (opm/models/discretization/common/tpfalinearizer.hh:116-379)
 *   - LinearSolverBackend   (opm/simulators/linalg/parallelbasebackend.hh:137-290)
 *   - NewtonMethod          (opm/models/nonlinear/newtonmethod.hh:112-438,
 *                            opm/models/blackoil/blackoilnewtonmethod.hh:54-448)
 * plus the time loop that calls them (opm/models/utils/simulator.hh:643-814,
 * opm/models/discretization/common/fvbaseproblem.hh:512-606).
 *
 * Conventions (drop-in layouts of the reference):
 *   - cells are numbered i + nx*(j + ny*k) (YaspGrid lexicographic, x fastest);
 *   - the Jacobian is block-CSR with sorted column indices per row and
 *     row-major 3x3 (or 2x2) blocks (Dune::BCRSMatrix<MatrixBlock>);
 *   - residual, solution and update vectors are cell-major with the
 *     equation / primary-variable index fastest (Dune::BlockVector);
 *   - black-oil equations: 0 oil, 1 water, 2 gas component;
 *     primary variables: 0 Sw, 1 po, 2 Sg or Rs (meaning byte 0 = Sg, 1 = Rs).
 *
 * Error convention: functions return 0 on success, 1 when a linear solve did
 * not converge, 2 for a recoverable numerical problem (the reference's
 * NumericalProblem, which makes the time loop halve the step), and a negative
 * value for fatal errors.  omodels_last_error() returns a thread-local message.
 *
 * Threading: one handle per HIP stream; calls on one handle are not
 * reentrant; different handles may be used from different threads.  Host
 * arrays are borrowed for the duration of a call only.
 */
#ifndef OMODELS_H
#define OMODELS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Problem description (host arrays, borrowed).  Same layout as ob_problem in
 * the oracle.  Replaces the reference's Problem/Vanguard pair for this path:
 * ReservoirProblem (tests/problems/reservoirproblem.hh:215-374) provides the
 * PVT tables and StructuredGridVanguard the grid. */
typedef struct omodels_blackoil_problem {
    int32_t nx, ny, nz, n_src;
    double dx, dy, dz;
    double gravity;          /* z component of gravity; z points up */
    const double* perm;      /* [3][N] kx, ky, kz [m^2] */
    const double* poro;      /* [N] */
    int32_t n_oil, n_gas;
    const double* oil_tab;   /* [4][n_oil]: p, Rs_sat, 1/Bo_sat, mu_o_sat */
    const double* gas_tab;   /* [3][n_gas]: p, 1/Bg, mu_g */
    double oil_comp, water_pref, water_comp, water_mu;
    double rho_ref[3];       /* water, oil, gas surface densities */
    const int32_t* src_cell;
    const int32_t* src_type; /* 0: water injector (mass rate), 1: BHP producer */
    const double* src_a;     /* injector: kg/s; producer: well index [m^3] */
    const double* src_b;     /* producer: bottom-hole pressure [Pa] */
} omodels_blackoil_problem;

/* Synthetic black-oil generator used by the benchmark configurations. */
typedef struct omodels_blackoil_spec {
    int32_t nx, ny, nz;
    int32_t field;           /* 0: i.i.d. log-normal kx; 1: channelised Gaussian field */
    double dx, dy, dz;
    double perm_mean_md;     /* field 0: geometric mean kx [mD] */
    double perm_sigma;       /* sigma of ln kx (within each facies for field 1) */
    double kz_ratio;         /* kz = kz_ratio * kx, ky = kx */
    double chan_perm_md, back_perm_md, chan_fraction, corr_len;  /* field 1 */
    double poro_min, poro_max;
    double p_top;            /* pressure of the top layer [Pa] */
    double sw_init;
    int32_t wells;           /* 0: corner injector/producer, 1: five-spot, 2: random pair */
    int32_t pad;
    double inj_rate;         /* total water injection [kg/s]; field 1 splits it over the
                                injector's layers by kh and places the five-spot injector
                                in the column of the largest kh near the centre */
    double prod_bhp;         /* producer bottom-hole pressure [Pa] */
    uint64_t seed;
    double rs_frac;          /* initial Rs as a fraction of the saturated value (0 or 1:
                                saturated oil, bubble point at the initial pressure) */
} omodels_blackoil_spec;

/* Fill the PVT tables of tests/problems/reservoirproblem.hh:223-327
 * (oil_tab [4][9], gas_tab [3][10]) and the fluid constants. */
int omodels_blackoil_default_fluid(double* oil_tab, int32_t* n_oil, double* gas_tab,
                                   int32_t* n_gas, double* consts /* [7] */);
/* Generate permeability, porosity, hydrostatic initial state and wells.
 * perm [3N], poro [N], pv [3N], meaning [N], sources up to src_cap. */
int omodels_blackoil_generate(const omodels_blackoil_spec* spec, double* perm, double* poro,
                              double* pv, uint8_t* meaning, int32_t src_cap, int32_t* n_src,
                              int32_t* src_cell, int32_t* src_type, double* src_a,
                              double* src_b);

/* ------------------------------------------------------------------------ */
/* Parameters, named after the reference's run-time parameters. */
typedef struct omodels_params {
    double newton_tol;          /* NewtonTolerance (newtonmethodparameters.hh:52) */
    int32_t newton_max_iter;    /* NewtonMaxIterations (:36) */
    int32_t newton_target_iter; /* NewtonTargetIterations (:45) */
    double newton_max_error;    /* NewtonMaxError (:32) */
    double lin_tol;             /* LinearSolverTolerance (linalgparameters.hh:63) */
    double lin_abs_tol;         /* LinearSolverAbsTolerance (<0: newton_tol/100) */
    double lin_max_error;       /* LinearSolverMaxError */
    int32_t lin_max_iter;       /* LinearSolverMaxIterations */
    int32_t precond;            /* 0: block ILU0 (PreconditionerWrapperILU), 1: block Jacobi,
                                   2: aggregation AMG (in place of ParallelAmgBackend,
                                   parallelamgbackend.hh:85-316; 3x3 grid systems, FP64) */
    double prec_relax;          /* PreconditionerRelaxation */
    int32_t prec_order;         /* PreconditionerOrder */
    int32_t max_dt_divisions;   /* MaxTimeStepDivisions (fvbaseparameters.hh:112) */
    double dp_max_rel;          /* DpMaxRel (blackoilnewtonmethodparameters.hh:34) */
    double ds_max;              /* DsMax */
    double privar_osc_thresh;   /* PriVarOscilationThreshold */
    double max_dt;              /* MaxTimeStepSize */
    double min_dt;              /* MinTimeStepSize */
    int32_t precision;          /* 0: FP64 solve; 1: FP32 Jacobian/ILU/inner BiCGStab
                                   inside FP64 flexible outer refinement; 2: FP32 solve
                                   (LinearSolverScalar = float, lens_immiscible_ecfv_ad.hh:53) */
    int32_t verbose;
    int32_t ilu_schedule;       /* 0: persistent tile-wavefront kernels, sweep form chosen from the
                                   tiles per tile level (or by the autotuner); 1: one launch per
                                   level; 2: tiles with streaming sweeps; 3: tiles with batched
                                   sweeps.  Ensembles and the server use tiles with batched sweeps
                                   unless 2 is given.  Results do not depend on the choice */
    int32_t autotune;           /* 1: time the ILU tile shapes (and the level schedule) on the first
                                   factorization of a grid size and keep the fastest; results do
                                   not depend on the choice */
    int32_t graph;              /* 1: each BiCGStab iteration is replayed from a captured hipGraph
                                   (one launch per iteration); results do not depend on it */
    int32_t spmv_in_sweep;      /* 1: with the streaming tiled ILU of one reservoir, the operator
                                   application after each preconditioner application runs inside
                                   the backward sweep, tile by tile as the tiles it reads finish
                                   (same results; no faster at 500k, where the sweeps have no idle
                                   workgroups); 0 (default): as a kernel of its own */
    int32_t ilu_order;          /* block ILU(0) ordering: 0 (default) the natural cell order of the
                                   reference's SeqILU; 1 two-colour tiles of 4x2x8 cells (colour-0
                                   tiles first, natural order inside a tile), whose sweeps run every
                                   tile of a colour at once -- a different preconditioner (more
                                   iterations, each far cheaper), for the FP32-inner and ensemble
                                   forms; 3x3 blocks, FP64 or FP32-inner solves */
    /* precond = 2: box aggregates of amg_agg cells per direction (the
     * reference's aggregateSizePerDim = 3, parallelamgbackend.hh:294-295),
     * Galerkin coarse operators, levels until one has at most
     * amg_coarsen_target cells (AmgCoarsenTarget, :72; here at most 32, the
     * coarsest level being solved by a dense inverse, where the reference
     * stops at 5000 with a sparse direct solver), a V-cycle with
     * amg_pre_smooth forward and amg_post_smooth backward block Gauss-Seidel
     * sweeps (SOR with relaxation 1, :272-276; red-black order) around
     * the coarse correction, prolongated with the factor amg_prolong_damp */
    int32_t amg_coarsen_target; /* 1..32 (default 32) */
    int32_t amg_pre_smooth;     /* >= 1 (default 2) */
    int32_t amg_post_smooth;    /* >= 0 (default 2) */
    double amg_prolong_damp;    /* default 1.6 */
    int32_t amg_agg[3];         /* >= 1 each (default 3, 3, 3) */
} omodels_params;

void omodels_default_params(omodels_params* prm);
const char* omodels_last_error(void);

/* ------------------------------------------------------------------------ */
/* Model: the device-resident state of one reservoir (FvBaseDiscretization's
 * solution history and storage cache, fvbasediscretization.hh:655-857). */
typedef struct omodels_model omodels_model;

omodels_model* omodels_model_create(const omodels_blackoil_problem* prob);
void omodels_model_destroy(omodels_model* m);
int omodels_model_num_cells(const omodels_model* m, int64_t* n, int64_t* nnzb);
/* solution(0) <- host; also solution(1) when both_levels != 0 */
int omodels_model_set_solution(omodels_model* m, const double* pv, const uint8_t* meaning,
                               int both_levels);
int omodels_model_get_solution(omodels_model* m, double* pv, uint8_t* meaning);
/* storage cache of time level 1 <- host (or NULL: recompute from solution(1)) */
int omodels_model_set_storage1(omodels_model* m, const double* storage1);
int omodels_model_get_storage1(omodels_model* m, double* storage1);

/* ------------------------------------------------------------------------ */
/* Linearizer -- replaces TpfaLinearizer (tpfalinearizer.hh:75-951). */
/* init(): createMatrix_ (tpfalinearizer.hh:414-525).  Exports the pattern
 * (row_ptr [N+1], col_idx [NNZB]), diagMatAddress_ as block offsets
 * (diag_off [N]) and per neighbour info in intersection order
 * (nb_ptr [N+1], nb_cell, nb_off = offset of A[J][I], nb_trans, nb_dzg). */
int omodels_linearizer_pattern(omodels_model* m, int32_t* row_ptr, int32_t* col_idx,
                               int32_t* diag_off, int32_t* nb_ptr, int32_t* nb_cell,
                               int32_t* nb_off, double* nb_trans, double* nb_dzg);
/* updateStoredTransmissibilities() (tpfalinearizer.hh:868-883) and the other
 * per-neighbour inputs createMatrix_ takes from the problem and the model
 * (:443-452), given by the caller in place of the structured grid's own:
 *   trans   [3][N] problem.transmissibility(I, J) of face I|I+e_axis at
 *           axis*N + I (entries of faces on the boundary are ignored);
 *   depth   [N] problem.dofCenterDepth(I), the cell centre's coordinate along
 *           gravity's axis: dZg = (depth_I - depth_J) * gravity; the grid's
 *           own is dz * (k + 0.5);
 *   volume  [N] model.dofTotalVolume(I) > 0: storefac = volume / dt (:776-778),
 *           the sources' volume (:834) and the Vin / Vex of the upwind
 *           decision when the pressure difference is zero;
 *   thpres  [3][N] problem.thresholdPressure(I, J) >= 0 at the face's index:
 *           a face whose |pressure difference| is below it carries no flow,
 *           above it the difference is reduced by it.
 * Each may be NULL for the grid's own (for thpres: none); a call replaces all
 * four.  The new values hold for every later linearization, FLOWS output and
 * simulate call of the model, and nb_trans / nb_dzg of
 * omodels_linearizer_pattern report them. */
int omodels_linearizer_update_discretization(omodels_model* m, const double* trans, const double* depth,
                                             const double* volume, const double* thpres);
/* Rock compressibility and compaction of the problem
 * (blackoilintensivequantities.hh:447-480, blackoilproblem.hh:131-181):
 *   compressibility [N] problem.rockCompressibility(I) >= 0 [1/Pa]: where > 0
 *           the porosity is multiplied by 1 + x + x^2/2, x = c (p_o - p_ref);
 *   ref_pressure [N] problem.rockReferencePressure(I) [Pa] (NULL: 1e5 Pa);
 *   table [3][n_table] (n_table in [2, 12], or 0 for none): oil pressure
 *           (increasing), rockCompPoroMultiplier and rockCompTransMultiplier
 *           (> 0), piecewise linear in p_o (linear beyond the ends); the
 *           porosity is multiplied by the first, and each face's
 *           transmissibility by the mean of its two cells' second
 *           (blackoillocalresidualtpfa.hh:372-381).
 * NULL compressibility and n_table 0 switch the rock terms off (the
 * reference's defaults: 0, 1e5 Pa, multipliers 1).  The storage cache,
 * linearizations, FLOWS, output fields and simulate calls include them. */
int omodels_model_set_rock(omodels_model* m, const double* compressibility, const double* ref_pressure,
                           int32_t n_table, const double* table);
/* Boundary conditions: the reference's boundary faces with a condition
 * (TpfaLinearizer::createMatrix_ collects them when the problem has
 * nonTrivialBoundaryConditions, tpfalinearizer.hh:476-497; linearize_ adds
 * their terms after the cell loop, :849-865; the term is
 * BlackOilLocalResidualTPFA::computeBoundaryFlux,
 * blackoillocalresidualtpfa.hh:507-620).  n faces; face[b] (0..5: -x, +x,
 * -y, +y, -z, +z) must be a face of cell[b] without a neighbour; type[b]:
 * 0 none (BCType::NONE), 1 rate (RATE: mass_rate[b][3], kg/(s m^2) in
 * equation order oil, water, gas, positive out of the cell), 2 free and 3
 * Dirichlet (FREE / DIRICHLET: the boundary's fluid state from the primary
 * variables state[b][3] with gas meaning meaning[b]).  trans[b]: the face's
 * transmissibility (NULL: the cell's half transmissibility A K d/|d|^2 of
 * transfluxmodule.hh:457-479); face_depth[b]: the face's depth (NULL: the
 * cell's depth for lateral faces, minus / plus half a cell for -z / +z);
 * the defaults are evaluated when the conditions are set.  The terms are
 * added at the end of linearize_domain, each cell's faces in face order;
 * blocks a caller adds with add_sparse_sources afterwards come after them.
 * n = 0 removes all.  The ensemble does not take boundary conditions. */
int omodels_model_set_boundary_conditions(omodels_model* m, int32_t n, const int32_t* cell, const int32_t* face,
                                          const int32_t* type, const double* mass_rate, const double* state,
                                          const uint8_t* meaning, const double* trans, const double* face_depth);
/* The boundary faces' entries of FLOWS (updateFlowsInfo, tpfalinearizer.hh:
 * 688-704: a face with a condition records its term, the values added to
 * the residual; the rest of the reference's boundary entries stay zero, and
 * FLORES records none) at solution(0): returns the number of faces set by
 * omodels_model_set_boundary_conditions and writes, for each in cell and
 * then face order, cell[b], face[b] and flows[b][3] (each may be NULL;
 * type-0 faces give zeros).  omodels_linearizer_flows holds the interior
 * faces. */
int omodels_linearizer_boundary_flows(omodels_model* m, int32_t* cell, int32_t* face, double* flows);
/* Auxiliary equations (BaseAuxiliaryModule, baseauxiliarymodule.hh:55-121):
 * n_dofs degrees of freedom after the N cells (rows N .. N + n_dofs - 1 of
 * the system, each with the model's three equations, setDofOffset) and the
 * pattern entries their module adds (addNeighbors, added by createMatrix_ at
 * tpfalinearizer.hh:503-507): block (row[k], col[k]) for k < n_links, rows
 * and columns in 0 .. N + n_dofs - 1; every auxiliary row needs its diagonal
 * block among them.  The exported pattern, Jacobian, residual, solver
 * vectors and Newton update then have N + n_dofs rows; the grid rows' blocks
 * keep their values, the auxiliary entries start from zero in each
 * linearization (resetSystem_); the auxiliary solution starts at zero
 * (applyInitial: set it with omodels_model_set_auxiliary_solution).  The
 * Newton error leaves the auxiliary rows out (newtonmethod.hh:583-585) and
 * their update is current - delta (:704-708).  The solver uses the
 * level-scheduled block ILU(0) or block Jacobi of the whole system (no AMG,
 * no two-colour ILU).  n_dofs = 0 returns to the grid's pattern.  FLOWS and
 * the ensemble do not see auxiliary DOFs. */
int omodels_model_set_auxiliary(omodels_model* m, int32_t n_dofs, int32_t n_links, const int32_t* row,
                                const int32_t* col);
/* solution(0) of the auxiliary DOFs, [n_dofs][3] (also solution(1) when
 * both_levels != 0) */
int omodels_model_set_auxiliary_solution(omodels_model* m, const double* x, int both_levels);
int omodels_model_get_auxiliary_solution(omodels_model* m, double* x);
/* The module's linearize(matrix, residual) (linearizeAuxiliaryEquations,
 * tpfalinearizer.hh:254-281): after linearize_domain, block k (row-major 3x3)
 * is added to the Jacobian's block (block_row[k], block_col[k]), which must
 * be in the pattern, and res[k][3] to the residual's row res_row[k], in the
 * given order.  Nothing is added when any entry is rejected. */
int omodels_linearizer_add_auxiliary(omodels_model* m, int32_t n_blocks, const int32_t* block_row,
                                     const int32_t* block_col, const double* blocks, int32_t n_res,
                                     const int32_t* res_row, const double* res);
/* The module inside omodels_simulate: fn(user, m, newton_iteration) is called
 * after every linearize_domain of the Newton loop (NewtonMethod::apply calls
 * linearizeAuxiliaryEquations_ there, newtonmethod.hh:287) and adds its
 * terms with omodels_linearizer_add_auxiliary; a nonzero return fails the
 * attempt like the reference's NumericalProblem (the step is then retried
 * with half the time step).  The auxiliary solution follows the time levels
 * with the cells'.  omodels_simulate with auxiliary DOFs needs it. */
typedef int (*omodels_aux_linearize_fn)(void* user, omodels_model* m, int newton_iteration);
int omodels_model_set_auxiliary_module(omodels_model* m, omodels_aux_linearize_fn fn, void* user);
/* Directional mobilities (BlackOilIntensiveQuantities::mobility(phaseIdx,
 * facedir), blackoilintensivequantities.hh:530-550, used by the face flux at
 * blackoillocalresidualtpfa.hh:341, :375-381).  The reference leaves the
 * directional relative permeabilities to the problem (updateRelperms,
 * multiphasebaseproblem.hh:257-261); here they are kr_dir = m_dir kr with the
 * caller's multipliers m [3][3][N] (axis x, y, z; phase water, oil, gas; cell),
 * finite and >= 0: the upwind cell's mobility along an axis is its mobility
 * times its multiplier.  NULL switches them off.  The assembly, FLOWS/FLORES
 * and the Newton runs of the model use them; the ensemble does not. */
int omodels_model_set_directional_mobility(omodels_model* m, const double* multipliers);
/* linearizeDomain() (tpfalinearizer.hh:189-245, linearize_ :711-866) at the
 * current solution(0); iteration_index 0 refreshes the storage cache of
 * time level 1 as recycleFirstIterationStorage() does (:786-812). */
int omodels_linearizer_linearize_domain(omodels_model* m, double dt, int iteration_index);
/* linearizeDomain(domain) (tpfalinearizer.hh:226-245) on the sub-domain of
 * the n distinct cells `cells`: resetSystem_(domain) zeroes their residual
 * rows and Jacobian rows (:370-379); linearize_ then evaluates only these
 * cells (:711-842): their residual rows and diagonal blocks, and each
 * cell's column blocks A[J][I] added into its neighbours' rows, which keep
 * their previous values when the neighbour lies outside the domain; the
 * storage cache of time level 1 is not refreshed (:793-803); the separate
 * sparse source terms and the boundary terms are added for all cells, as
 * the loops after the cell loop do (:843-865).  All cells: the same as
 * linearize_domain.  Not with diffusion or dispersion. */
int omodels_linearizer_linearize_subdomain(omodels_model* m, double dt, int iteration_index, int32_t n,
                                           const int32_t* cells);
/* jacobian() / residual() host views (Dune layouts) */
int omodels_linearizer_get_jacobian(omodels_model* m, double* values);
int omodels_linearizer_get_residual(omodels_model* m, double* residual);
/* eraseMatrix(): release the device Jacobian (re-created on next linearize) */
int omodels_linearizer_erase_matrix(omodels_model* m);
/* SeparateSparseSourceTerms (tpfalinearizer.hh:56, default off): when on,
 * linearize_domain leaves the problem's source terms out of the cell loop
 * (computeSourceDense instead of computeSource, :829-833) and adds them after
 * it as sparse blocks, as wellModel().addReservoirSourceTerms does (:842-846);
 * the problem's source list plays the well model.  Either way each cell's
 * source term is the last addition to its row.  The cell loop adds a cell's
 * one rate (problem.source: the sum of the cell's sources in the list's
 * order, times -V); the sparse path adds each source's block in turn.  Where
 * no cell has two sources both settings give the same bits.  Further terms
 * (an external well model's) go in through add_sparse_sources after
 * linearize_domain. */
int omodels_linearizer_set_separate_sparse_sources(omodels_model* m, int enable);
/* The problem's source terms at the current solution(0) as sparse blocks:
 * cells [n_src], residual contributions [n_src][3] (already times -V, the
 * linearizer's units) and their derivatives w.r.t. the cell's primaries
 * [n_src][3][3] row-major; returns n_src.  NULL arrays are not written. */
int omodels_linearizer_source_blocks(omodels_model* m, int32_t* cells, double* residual, double* diag);
/* updateFlowsInfo (tpfalinearizer.hh:647-705) at the current solution(0), for
 * the FLOWS and FLORES output: per cell and per neighbour in intersection order
 * (the order of nb_ptr / nb_cell of omodels_linearizer_pattern, nnzb - N
 * entries), flows = the values of the face's component fluxes times the face
 * area as the residual takes them, flores = each phase's darcyFlux.value() *
 * faceArea at its component index (blackoillocalresidualtpfa.hh:376-385);
 * both [nnzb - N][3], either may be NULL.  The face direction of an entry is
 * its position among the cell's existing neighbours -x, +x, -y, +y, -z, +z
 * (the reference's dirId 0..5, ecfvstencil.hh:209-218). */
int omodels_linearizer_flows(omodels_model* m, double* flows, double* flores);

/* Molecular diffusion in the oil phase: the model with enableDiffusion, whose
 * face term is BlackOilDiffusionModule::addDiffusiveFlux
 * (blackoildiffusionmodule.hh:168-238, added in
 * blackoillocalresidualtpfa.hh:469-479).  The oil component and the dissolved
 * gas move along the Rs difference of the two cells of a face, weighted by the
 * mean So/Bo, in surface-volume units as the reference adds them.  Water has
 * no term (no gas in water) and the gas phase none without vaporized oil. */
typedef struct omodels_diffusion {
    double oil_in_oil;              /* coefficient of the oil component in the oil phase [m^2/s] */
    double gas_in_oil;              /* coefficient of the gas component in the oil phase [m^2/s] */
    double mw_oil, mw_gas;          /* molar masses; both > 0: mole-fraction form (DIFFC's
                                     * diffMoleFraction, :118-121), else mass fractions */
    const double* face_diffusivity; /* the problem's diffusivity(I, J) (tpfalinearizer.hh:462-464),
                                     * [3][N] with the face I | I + e_axis at axis*N + I;
                                     * NULL: the harmonic mean of A phi |n.d| / |d|^2 */
} omodels_diffusion;
/* Switch diffusion on for the following linearizations and flows (d != NULL)
 * or off (d == NULL).  No reference counterpart as a call: the reference
 * fixes enableDiffusion at compile time and reads the coefficients from the
 * deck.  The ensemble path has no diffusion. */
int omodels_model_set_diffusion(omodels_model* m, const omodels_diffusion* d);

/* Dispersion in the oil phase: the model with DISPERC, whose face term is
 * BlackOilDispersionModule::addDispersiveFlux (blackoildispersionmodule.hh:176-254,
 * added in blackoillocalresidualtpfa.hh:481-493): the oil component and the
 * dissolved gas move along the Rs difference, weighted by the mean 1/Bo and
 * the mean of the two cells' oil velocity.  A cell's velocity is the largest
 * |oil Darcy flux per area| over its faces in the previous linearization
 * (tpfalinearizer.hh:759-761 with blackoildispersionmodule.hh:355-383); it is
 * zero before the first (createFlows_, tpfalinearizer.hh:571-619). */
typedef struct omodels_dispersion {
    const double* face_dispersivity; /* the problem's dispersivity(I, J) (tpfalinearizer.hh:465-467),
                                      * [3][N] like face_diffusivity, or NULL */
    const double* cell_dispersion;   /* [N] dispersion coefficients [m], used when the face array is
                                      * NULL: harmonic mean of A E |n.d| / |d|^2 */
} omodels_dispersion;
/* Switch dispersion on (d != NULL: the cells' velocities start at zero) or off
 * (d == NULL).  No reference counterpart as a call (DISPERC in the deck). */
int omodels_model_set_dispersion(omodels_model* m, const omodels_dispersion* d);
/* The cells' oil velocities [N] the next linearization will use (the
 * normVelocityCell of BlackOilDispersionIntensiveQuantities); dispersion on. */
int omodels_model_get_cell_velocity(omodels_model* m, double* velocity);

/* ------------------------------------------------------------------------ */
/* Output and restart files. */
/* Output quantities of solution(0), [count][N] field after field, named as
 * the reference's VTK modules name them: pressure_, saturation_, density_ and
 * mobility_<water|oil|gas> and porosity (VtkMultiPhaseModule,
 * vtkmultiphasemodule.hh:306-323), R_s, B_w, B_o, B_g and "primary vars
 * meaning gas" (VtkBlackOilModule, vtkblackoilmodule.hh:276-301). */
int omodels_output_field_count(void);
const char* omodels_output_field_name(int32_t i);   /* NULL out of range */
int omodels_output_fields(omodels_model* m, double* out);
/* VTK unstructured-grid file (.vtu, ASCII) of cell fields [nfields][N] on the
 * hexahedra of an nx x ny x nz grid with spacings dx, dy, dz (the files
 * VtkMultiWriter writes, vtkmultiwriter.hh); host only. */
int omodels_vtk_write(const char* path, int32_t nx, int32_t ny, int32_t nz, double dx, double dy, double dz,
                      int32_t nfields, const char* const* names, const double* fields);
/* the output quantities of solution(0) as a .vtu file */
int omodels_model_write_vtk(omodels_model* m, const char* path);
/* Restart files in the reference's text format (Restart, restart.hh:40-280;
 * Simulator::serialize, simulator.hh:928-940; BlackOilModel::serializeEntity,
 * blackoilmodel.hh:468-492): grid cookie, simulator section (time and time
 * step index), one line per cell with the primaries, the gas / water /
 * pressure meanings and the PVT region; host only.  Reading checks the grid
 * cookie and every section and rejects meanings this model does not have. */
int omodels_restart_write(const char* path, int32_t nx, int32_t ny, int32_t nz, const double* pv,
                          const uint8_t* meaning, double time, int32_t time_step_idx);
int omodels_restart_read(const char* path, int32_t nx, int32_t ny, int32_t nz, double* pv, uint8_t* meaning,
                         double* time, int32_t* time_step_idx);
/* solution(0) to a restart file; from one into solution(0) and solution(1) */
int omodels_model_write_restart(omodels_model* m, const char* path, double time, int32_t time_step_idx);
int omodels_model_read_restart(omodels_model* m, const char* path, double* time, int32_t* time_step_idx);
/* addReservoirSourceTerms(residual, diagMatAddress) (tpfalinearizer.hh:843-846):
 * residual[cell] += r and A[cell][cell] += d for every entry, in the given
 * order (a cell may appear more than once). */
int omodels_linearizer_add_sparse_sources(omodels_model* m, int32_t n, const int32_t* cells,
                                          const double* residual, const double* diag);

/* ------------------------------------------------------------------------ */
/* Linear solver backend -- replaces ParallelBiCGStabSolverBackend
 * (parallelbicgstabbackend.hh:91-172 over parallelbasebackend.hh:107-386). */
typedef struct omodels_solver omodels_solver;
omodels_solver* omodels_solver_create(omodels_model* m, const omodels_params* prm);
void omodels_solver_destroy(omodels_solver* s);
/* prepare(J, r) + setResidual(r) with the model's linearizer output, or with
 * host arrays when values/residual are non-NULL (BCRS layout of the model). */
int omodels_solver_prepare(omodels_solver* s, const double* values, const double* residual);
int omodels_solver_get_residual(omodels_solver* s, double* residual);
/* setMatrix + solve(x) (:252-284); x host [N*bs] may be NULL (kept on device) */
int omodels_solver_solve(omodels_solver* s, double* x);
int omodels_solver_iterations(const omodels_solver* s);
/* Scalar product of the backend's BiCGStab, OverlappingScalarProduct::dot
 * (overlappingscalarproduct.hh:58-70) for one process: *out = sum x[i]*y[i]
 * over host arrays of m < 2^28 entries.  The sum is exact and rounded once to
 * the nearest double (ties to even), so it does not depend on the order of
 * the additions; the reference's sequential loop approximates this value. */
int omodels_scalar_product(int64_t m, const double* x, const double* y, double* out);
/* The same scalar product without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * x_dev, y_dev: device arrays of m < 2^28 doubles; out_dev: one device
 * double; stream: the caller's hipStream_t (NULL: the null stream).  The work
 * is only enqueued: the result is in out_dev once the stream has reached it.
 * The partial sums live in a static device array of 1.3 KiB per GPU, shared
 * by all calls on that GPU: calls that overlap on one GPU race. */
int omodels_scalar_product_noalloc(int64_t m, const double* x_dev, const double* y_dev, double* out_dev,
                                   void* stream);
/* Block ILU0 factor of the prepared matrix (diagonal blocks inverted), for
 * parity checks of PreconditionerWrapperILU (istlpreconditionerwrappers.hh:151-190). */
int omodels_solver_get_ilu(omodels_solver* s, double* lu_values);
/* Preconditioner apply v = M^-1 d on host arrays. */
int omodels_solver_apply_precond(omodels_solver* s, const double* d, double* v);
/* The aggregation AMG (precond = 2) of the prepared matrix, for parity
 * checks: the number of levels (dims[3 * l + q], q = x, y, z, for the first
 * `cap` levels), and level l's matrix in the structured layout [7][N_l][9]
 * (slots -z, -y, -x, I, +x, +y, +z; absent neighbours' blocks zero). */
int omodels_solver_amg_levels(omodels_solver* s, int32_t* dims, int cap);
int omodels_solver_get_amg_level(omodels_solver* s, int level, double* values);

/* ------------------------------------------------------------------------ */
/* Newton method + time loop -- replaces BlackOilNewtonMethod::apply and
 * FvBaseProblem::timeIntegration / Simulator::run for n_steps time steps. */
typedef struct omodels_iter_log {
    int32_t step, attempt, iter, lin_iters;
    int32_t num_switched, lin_converged;
    double error, dt;
} omodels_iter_log;

typedef struct omodels_run_stats {
    int32_t newton_iters, linearizations, linear_iters, failed_attempts;
    double t_linearize, t_solve, t_update, t_total;   /* seconds, host wall clock */
} omodels_run_stats;

/* Apply the Newton update of the black-oil model to the device state
 * (BlackOilNewtonMethod::updatePrimaryVariables_, blackoilnewtonmethod.hh:226-427)
 * with a host update vector; used for parity checks of the update kernel. */
int omodels_newton_update(omodels_model* m, const omodels_params* prm, const double* delta,
                          int32_t* num_switched);
int omodels_model_get_switched(omodels_model* m, uint8_t* was_switched);

int omodels_simulate(omodels_model* m, const omodels_params* prm, int n_steps, double dt0,
                     double* pv_hist, double* dt_hist, omodels_iter_log* log, int log_cap,
                     int32_t* log_len, omodels_run_stats* stats);

/* ------------------------------------------------------------------------ */
/* Ensemble engine and request loop.  No counterpart in the reference (which
 * runs one reservoir per process); each member follows the reference's
 * Newton/time-step control (newtonmethod.hh:228-465, fvbaseproblem.hh:512-606)
 * while the kernels of all members run in the same launches.  Requests are
 * admitted in arrival order when their device memory fits hbm_budget_bytes
 * and at most max_batch run at once.  The batch advances in rounds of one
 * Newton iteration per running member; each member starts and ends its time
 * steps on its own (no member waits for another's step), a finished request
 * leaves and an arrived one joins between rounds.  All members share one
 * fluid. */
typedef struct omodels_server omodels_server;
omodels_server* omodels_server_create(const omodels_params* prm, double hbm_budget_bytes, int max_batch);
void omodels_server_destroy(omodels_server* s);
/* copies the problem; returns the request id (>= 0) */
int omodels_server_submit(omodels_server* s, const omodels_blackoil_problem* p, const double* pv,
                          const uint8_t* meaning, int n_steps, double dt0, double arrival_s);
/* runs until every submitted request is done (arrival times are seconds
 * after the start of this call) */
int omodels_server_run(omodels_server* s);
/* final state, times [arrival, admission, completion] (s), per-step dt and
 * statistics of one request; returns its status (0 ok, 3 time step failed) */
int omodels_server_result(omodels_server* s, int id, double* pv, uint8_t* meaning, double* times,
                          double* dt_hist, omodels_run_stats* stats);
int omodels_server_log(omodels_server* s, int id, omodels_iter_log* log, int cap, int32_t* len);
/* wall time of the last run, batched Newton rounds, most members at once */
int omodels_server_summary(omodels_server* s, double* wall_s, int64_t* batch_rounds, int32_t* max_running);
/* Test hook for the failure handling of the tiled ILU's bounded dependency
 * waits (ilu_order 0): sets the error flag that a wait which ran into its
 * bound sets, so the waits of the next batched solve give up at once; every
 * member of that solve then fails it as a NumericalProblem (its time step is
 * halved and retried) and the flag and exchange slots are reset.  No
 * counterpart in the reference (whose sweeps are sequential). */
int omodels_server_inject_ilu_wait_error(omodels_server* s);
int omodels_server_set_profiling(omodels_server* s, int enable);
int omodels_server_kernel_times(omodels_server* s, double* ms, int64_t* counts, int n);

/* ------------------------------------------------------------------------ */
/* Kernel timing with HIP events recorded on the model's stream around each
 * kernel class: 0 assembly, 1 SpMV, 2 preconditioner apply (ILU0 sweep pair),
 * 3 ILU0 factorization, 4 other.  set_profiling(enable): 0 off, k >= 1 times
 * every k-th launch of each class (each event record idles the stream for a
 * few microseconds); kernel_times reports per class the launch count and the
 * timed launches' average duration times that count.  The reference's
 * counterpart is its phase timers (newtonmethod.hh:240-262, timer.hh:48-215). */
int omodels_model_set_profiling(omodels_model* m, int enable);
int omodels_model_kernel_times(omodels_model* m, double* ms, int64_t* counts, int n);

/* ------------------------------------------------------------------------ */
/* Two-phase immiscible element path (config 1: the lens problem).
 * Replaces, for ImmiscibleModel + EcfvDiscretization + AutoDiffLocalLinearizer
 * (tests/lens_immiscible_ecfv_ad.hh:37-55): FvBaseLinearizer::linearize_ /
 * linearizeElement_ (fvbaselinearizer.hh:510-626), FvBaseAdLocalLinearizer
 * (fvbaseadlocallinearizer.hh:173-292), FvBaseLocalResidual::eval
 * (fvbaselocalresidual.hh:158-626), the problem of tests/problems/lensproblem.hh:227-627
 * and the Newton / time loop (newtonmethod.hh:228-465, fvbaseproblem.hh:512-606,
 * simulator.hh:481-512,697-810).  2-D structured grid, cell I = i + nx*j,
 * primary variables (p_w [Pa], S_w) per cell, 2x2 row-major blocks. */
typedef struct omodels_lens_problem {
    int32_t nx, ny;
    double lx, ly;                                   /* domain size [m] */
    double lens_x0, lens_y0, lens_x1, lens_y1;       /* lens box [m] (lensproblem.hh:130-148) */
    double k_lens, k_outer;                          /* intrinsic permeability [m^2] */
    double porosity;
    double swr_lens, snr_lens, alpha_lens, n_lens;   /* regularized Van Genuchten, lens */
    double swr_outer, snr_outer, alpha_outer, n_outer;
    double rho_w, rho_n, mu_w, mu_n;                 /* incompressible liquids */
    double gravity;                                  /* y component [m/s^2] */
    double inflow_rate;                              /* DNAPL inflow on the top strip [kg/(m^2 s)] */
    double p_atm;                                    /* pressure at the top of the free-flow sides [Pa] */
    double eps;                                      /* geometric tolerance [m] */
} omodels_lens_problem;

typedef struct omodels_lens omodels_lens;

/* the defaults of LensProblem (lensproblem.hh:254-262,301-320,370-373) */
void omodels_lens_default_problem(omodels_lens_problem* p);
omodels_lens* omodels_lens_create(const omodels_lens_problem* p);
void omodels_lens_destroy(omodels_lens* m);
int omodels_lens_num_cells(const omodels_lens* m, int64_t* n, int64_t* nnzb);
/* createMatrix_ (fvbaselinearizer.hh:382-420): sorted columns; diag optional */
int omodels_lens_pattern(omodels_lens* m, int32_t* row_ptr, int32_t* col_idx, int32_t* diag);
/* LensProblem::initial (lensproblem.hh:550-582): water-saturated, hydrostatic */
int omodels_lens_initial(omodels_lens* m, double* pv);
int omodels_lens_set_solution(omodels_lens* m, const double* pv, int both_levels);
int omodels_lens_get_solution(omodels_lens* m, double* pv);
int omodels_lens_set_storage1(omodels_lens* m, const double* storage1);
/* linearize() at the current solution with the storage cache of the
 * previous level (computed at iteration 0 from the current solution) */
int omodels_lens_linearize(omodels_lens* m, double dt, int iteration_index);
int omodels_lens_get_jacobian(omodels_lens* m, double* values);
int omodels_lens_get_residual(omodels_lens* m, double* residual);
/* Simulator::run until t_end from dt0; t_hist / dt_hist have max_steps entries */
int omodels_lens_simulate(omodels_lens* m, const omodels_params* prm, double t_end, double dt0,
                          int max_steps, double* t_hist, double* dt_hist, int32_t* n_steps,
                          omodels_iter_log* log, int log_cap, int32_t* log_len,
                          omodels_run_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
