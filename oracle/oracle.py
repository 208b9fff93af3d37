"""ctypes binding of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the parity reference and CPU baseline.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


class ObProblem(C.Structure):
    """ob_problem (oracle.h): omodels_blackoil_problem's layout plus the
    caller-given discretization (NULL: the structured grid's own)."""
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("n_src", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("gravity", C.c_double),
        ("perm", _dp), ("poro", _dp),
        ("n_oil", C.c_int32), ("n_gas", C.c_int32),
        ("oil_tab", _dp), ("gas_tab", _dp),
        ("oil_comp", C.c_double), ("water_pref", C.c_double), ("water_comp", C.c_double),
        ("water_mu", C.c_double), ("rho_ref", C.c_double * 3),
        ("src_cell", _ip), ("src_type", _ip), ("src_a", _dp), ("src_b", _dp),
        ("face_trans", _dp), ("depth", _dp), ("volume", _dp), ("thpres", _dp),
        ("rock_comp", _dp), ("rock_pref", _dp), ("n_rock", C.c_int32), ("rock_tab", _dp), ("dir_mob", _dp),
        ("n_bc", C.c_int32), ("bc_cell", _ip), ("bc_face", _ip), ("bc_type", _ip), ("bc_rate", _dp),
        ("bc_state", _dp), ("bc_meaning", _u8p), ("bc_trans", _dp), ("bc_zface", _dp),
        ("n_aux", C.c_int32), ("n_aux_links", C.c_int32), ("aux_row", _ip), ("aux_col", _ip),
    ]


class ObDiffusion(C.Structure):
    """ob_diffusion (oracle.h), same layout as omodels_diffusion."""
    _fields_ = [("oil_in_oil", C.c_double), ("gas_in_oil", C.c_double), ("mw_oil", C.c_double),
                ("mw_gas", C.c_double), ("face_diffusivity", _dp)]


def diffusion(oil_in_oil, gas_in_oil, mw_oil=0.0, mw_gas=0.0, face_diffusivity=None):
    """ObDiffusion; keeps a reference to the face array it points at."""
    d = ObDiffusion(oil_in_oil, gas_in_oil, mw_oil, mw_gas, None)
    if face_diffusivity is not None:
        fd = np.ascontiguousarray(face_diffusivity, dtype=np.float64)
        d.face_diffusivity = fd.ctypes.data_as(_dp)
        d._keep = fd
    return d


class ObDispersion(C.Structure):
    """ob_dispersion (oracle.h), same layout as omodels_dispersion."""
    _fields_ = [("face_dispersivity", _dp), ("cell_dispersion", _dp)]


def dispersion(face_dispersivity=None, cell_dispersion=None):
    d = ObDispersion(None, None)
    d._keep = []
    for name, a in (("face_dispersivity", face_dispersivity), ("cell_dispersion", cell_dispersion)):
        if a is not None:
            a = np.ascontiguousarray(a, dtype=np.float64)
            setattr(d, name, a.ctypes.data_as(_dp))
            d._keep.append(a)
    return d


class ObParams(C.Structure):
    _fields_ = [
        ("newton_tol", C.c_double), ("newton_max_iter", C.c_int32), ("newton_target_iter", C.c_int32),
        ("newton_max_error", C.c_double), ("lin_tol", C.c_double), ("lin_abs_tol", C.c_double),
        ("lin_max_error", C.c_double), ("lin_max_iter", C.c_int32), ("precond", C.c_int32),
        ("prec_relax", C.c_double), ("prec_order", C.c_int32), ("max_dt_divisions", C.c_int32),
        ("dp_max_rel", C.c_double), ("ds_max", C.c_double), ("privar_osc_thresh", C.c_double),
        ("max_dt", C.c_double), ("min_dt", C.c_double), ("solver_float", C.c_int32), ("verbose", C.c_int32),
        ("ilu_order", C.c_int32), ("amg_coarsen_target", C.c_int32), ("amg_pre_smooth", C.c_int32),
        ("amg_post_smooth", C.c_int32), ("amg_prolong_damp", C.c_double), ("amg_agg", C.c_int32 * 3),
    ]


class ObIterLog(C.Structure):
    _fields_ = [("step", C.c_int32), ("attempt", C.c_int32), ("iter", C.c_int32), ("lin_iters", C.c_int32),
                ("num_switched", C.c_int32), ("lin_converged", C.c_int32), ("error", C.c_double),
                ("dt", C.c_double)]


class OlProblem(C.Structure):
    """Same layout as ol_problem (oracle.h) and omodels_lens_problem."""
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("lx", C.c_double), ("ly", C.c_double),
        ("lens_x0", C.c_double), ("lens_y0", C.c_double), ("lens_x1", C.c_double), ("lens_y1", C.c_double),
        ("k_lens", C.c_double), ("k_outer", C.c_double), ("porosity", C.c_double),
        ("swr_lens", C.c_double), ("snr_lens", C.c_double), ("alpha_lens", C.c_double), ("n_lens", C.c_double),
        ("swr_outer", C.c_double), ("snr_outer", C.c_double), ("alpha_outer", C.c_double),
        ("n_outer", C.c_double),
        ("rho_w", C.c_double), ("rho_n", C.c_double), ("mu_w", C.c_double), ("mu_n", C.c_double),
        ("gravity", C.c_double), ("inflow_rate", C.c_double), ("p_atm", C.c_double), ("eps", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        L.ob_default_params.argtypes = [C.POINTER(ObParams)]
        L.ob_pattern_size.argtypes = [C.POINTER(ObProblem), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.ob_build_pattern.argtypes = [C.POINTER(ObProblem), _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _ip]
        L.ob_storage.argtypes = [C.POINTER(ObProblem), _dp, _u8p, _dp]
        L.ob_linearize.argtypes = [C.POINTER(ObProblem), _dp, _u8p, _dp, C.c_double, _dp, _dp, C.c_int]
        L.ob_linearize_timed.argtypes = [C.POINTER(ObProblem), _dp, _u8p, _dp, C.c_double, C.c_int, C.c_int, _dp]
        L.ob_source_blocks.argtypes = [C.POINTER(ObProblem), _dp, _u8p, _ip, _dp, _dp]
        L.ob_flows.argtypes = [C.POINTER(ObProblem), _dp, _u8p, _dp, _dp]
        L.ob_intensive.argtypes = [C.POINTER(ObProblem), _dp, C.c_int, C.c_int, _dp]
        L.ob_ilu0_factor.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp]
        L.ob_ilu0_apply.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_double]
        L.ob_spmv.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp]
        L.ob_dot.argtypes = [C.c_int64, _dp, _dp]
        L.ob_colored_ilu.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_double]
        L.ob_bicgstab_grid.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, C.POINTER(ObParams), _ip]
        L.ob_jacobi_apply.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_int, C.c_double]
        L.ob_amg_dims.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(ObParams), _ip, C.c_int]
        L.ob_amg_level_matrix.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, C.POINTER(ObParams), C.c_int,
                                          _dp]
        L.ob_amg_apply.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, C.POINTER(ObParams), _dp, _dp]
        L.ob_gs_sweep.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, C.c_int]
        L.ob_dense_inverse.argtypes = [C.c_int, _dp]
        L.ob_dot.restype = C.c_double
        L.ob_simulate_aux.argtypes = [C.POINTER(ObProblem), _dp, _u8p, _dp, OB_AUX_FN, C.c_void_p, C.c_int,
                                      C.c_double, C.POINTER(ObParams), _dp, _dp, C.POINTER(ObIterLog), C.c_int,
                                      _ip, C.c_int]
        L.ob_boundary_flows.argtypes = [C.POINTER(ObProblem), _dp, _u8p, _ip, _ip, _dp]
        L.ob_linearize_domain.argtypes = [C.POINTER(ObProblem), _dp, _u8p, _dp, C.c_double, C.c_int32, _ip, _dp,
                                          _dp]
        L.ob_add_auxiliary.argtypes = [C.POINTER(ObProblem), C.c_int32, _ip, _ip, _dp, C.c_int32, _ip, _dp, _dp,
                                       _dp]
        L.ob_bicgstab.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, C.POINTER(ObParams), _ip]
        L.ob_update_primary.argtypes = [C.POINTER(ObProblem), _dp, _u8p, _dp, _dp, _u8p, _u8p,
                                        C.POINTER(ObParams), _ip]
        L.ob_simulate.argtypes = [C.POINTER(ObProblem), _dp, _u8p, C.c_int, C.c_double, C.POINTER(ObParams),
                                  _dp, _dp, C.POINTER(ObIterLog), C.c_int, _ip, _dp, C.c_int]
        _dif = C.POINTER(ObDiffusion)
        _dsp = C.POINTER(ObDispersion)
        L.ob_linearize_ext.argtypes = [C.POINTER(ObProblem), _dif, _dsp, _dp, _dp, _dp, _u8p, _dp, C.c_double, _dp,
                                       _dp, C.c_int]
        L.ob_flows_ext.argtypes = [C.POINTER(ObProblem), _dif, _dsp, _dp, _dp, _u8p, _dp, _dp]
        L.ob_simulate_ext.argtypes = [C.POINTER(ObProblem), _dif, _dsp, _dp, _u8p, C.c_int, C.c_double,
                                      C.POINTER(ObParams), _dp, _dp, C.POINTER(ObIterLog), C.c_int, _ip, _dp,
                                      C.c_int]
        L.ol_default_problem.argtypes = [C.POINTER(OlProblem)]
        L.ol_num_cells.argtypes = [C.POINTER(OlProblem), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.ol_pattern.argtypes = [C.POINTER(OlProblem), _ip, _ip, _ip]
        L.ol_initial.argtypes = [C.POINTER(OlProblem), _dp]
        L.ol_storage.argtypes = [C.POINTER(OlProblem), _dp, _dp]
        L.ol_linearize.argtypes = [C.POINTER(OlProblem), _dp, _dp, C.c_double, _dp, _dp]
        L.ol_material.argtypes = [C.POINTER(OlProblem), C.c_int, C.c_double, _dp]
        L.ol_simulate.argtypes = [C.POINTER(OlProblem), _dp, C.c_double, C.c_double, C.POINTER(ObParams),
                                  C.c_int, _dp, _dp, _ip, C.POINTER(ObIterLog), C.c_int, _ip]
        _lib = L
    return _lib


def P(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def params(**kw):
    p = ObParams()
    lib().ob_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "amg_agg":
            for q in range(3):
                p.amg_agg[q] = int(v[q])
        else:
            setattr(p, k, v)
    return p


def problem_from_case(case):
    """ob_problem view of a BlackOilCase built by the product's generator."""
    return case.descriptor(ObProblem)


def with_discretization(prob, trans=None, depth=None, volume=None, thpres=None):
    """prob with the caller-given transmissibilities / thresholds ([3][N] at
    the face's index) and depths / volumes ([N]); None for the grid's own.
    The arrays are kept alive by the returned problem."""
    N = prob.nx * prob.ny * prob.nz
    keep = []
    for name, a, n in (("face_trans", trans, 3 * N), ("depth", depth, N), ("volume", volume, N),
                       ("thpres", thpres, 3 * N)):
        if a is None:
            setattr(prob, name, None)
            continue
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        assert a.size == n, (name, a.size, n)
        keep.append(a)
        setattr(prob, name, P(a))
    prob._geo_keep = keep
    return prob


def with_rock(prob, comp=None, pref=None, table=None):
    """prob with rock compressibilities / reference pressures ([N]) and the
    multiplier table ([3][n]: p, pore-volume, transmissibility multipliers);
    None for none.  The arrays are kept alive by the returned problem."""
    keep = []
    for name, a in (("rock_comp", comp), ("rock_pref", pref), ("rock_tab", table)):
        if a is None:
            setattr(prob, name, None)
            continue
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        keep.append(a)
        setattr(prob, name, P(a))
    prob.n_rock = 0 if table is None else len(np.ravel(table)) // 3
    prob._rock_keep = keep
    return prob


def with_directional_mobility(prob, mult):
    """prob with directional relative permeability multipliers [3][3][N]
    (axis, phase water/oil/gas, cell); None for none"""
    if mult is None:
        prob.dir_mob = None
        prob._dm_keep = None
        return prob
    a = np.ascontiguousarray(mult, dtype=np.float64).ravel()
    prob.dir_mob = P(a)
    prob._dm_keep = a
    return prob


def with_boundary(prob, cell=None, face=None, type=None, rate=None, state=None, meaning=None, trans=None,
                  zface=None):
    """prob with boundary conditions (see ob_problem); cell None for none"""
    if cell is None:
        prob.n_bc = 0
        prob._bc_keep = None
        return prob
    keep = {}
    n = len(cell)
    for name, a, dt, pt in (("bc_cell", cell, np.int32, _ip), ("bc_face", face, np.int32, _ip),
                            ("bc_type", type, np.int32, _ip), ("bc_rate", rate, np.float64, _dp),
                            ("bc_state", state, np.float64, _dp), ("bc_meaning", meaning, np.uint8, _u8p),
                            ("bc_trans", trans, np.float64, _dp), ("bc_zface", zface, np.float64, _dp)):
        if a is None:
            setattr(prob, name, None)
            continue
        arr = np.ascontiguousarray(a, dtype=dt).ravel()
        keep[name] = arr
        setattr(prob, name, P(arr, pt))
    prob.n_bc = n
    prob._bc_keep = keep
    return prob


def with_auxiliary(prob, n_aux=0, rows=None, cols=None):
    """prob with n_aux auxiliary degrees of freedom after the cells and the
    pattern entries (rows[k], cols[k]) their module adds; 0 for none"""
    if n_aux == 0:
        prob.n_aux = prob.n_aux_links = 0
        prob.aux_row = prob.aux_col = None
        prob._aux_keep = None
        return prob
    r = np.ascontiguousarray(rows, dtype=np.int32).ravel()
    c = np.ascontiguousarray(cols, dtype=np.int32).ravel()
    prob.n_aux, prob.n_aux_links = int(n_aux), r.size
    prob.aux_row, prob.aux_col = P(r, _ip), P(c, _ip)
    prob._aux_keep = (r, c)
    return prob


def rows(prob):
    """rows of the system: the cells and the auxiliary degrees of freedom"""
    return prob.nx * prob.ny * prob.nz + prob.n_aux


def add_auxiliary(prob, J, r, brow=(), bcol=(), blocks=None, rrow=(), rvals=None):
    """(J, r) with the auxiliary module's blocks and residual entries added in
    order (ob_add_auxiliary)"""
    J = np.array(J, dtype=np.float64)
    r = np.array(r, dtype=np.float64)
    br = np.ascontiguousarray(brow, dtype=np.int32).ravel()
    bc = np.ascontiguousarray(bcol, dtype=np.int32).ravel()
    bl = np.ascontiguousarray(np.zeros(0) if blocks is None else blocks, dtype=np.float64).ravel()
    rr = np.ascontiguousarray(rrow, dtype=np.int32).ravel()
    rv = np.ascontiguousarray(np.zeros(0) if rvals is None else rvals, dtype=np.float64).ravel()
    assert bl.size == 9 * br.size and rv.size == 3 * rr.size
    rc = lib().ob_add_auxiliary(C.byref(prob), br.size, P(br, _ip), P(bc, _ip), P(bl), rr.size, P(rr, _ip), P(rv),
                                P(J), P(r))
    assert rc == 0
    return J, r


def pattern(prob):
    nz = C.c_int64()
    f2 = C.c_int64()
    assert lib().ob_pattern_size(C.byref(prob), C.byref(nz), C.byref(f2)) == 0
    N = rows(prob)
    rp = np.zeros(N + 1, np.int32)
    ci = np.zeros(nz.value, np.int32)
    dg = np.zeros(N, np.int32)
    nbp = np.zeros(prob.nx * prob.ny * prob.nz + 1, np.int32)
    nbc = np.zeros(f2.value, np.int32)
    nbo = np.zeros(f2.value, np.int32)
    nbt = np.zeros(f2.value)
    nba = np.zeros(f2.value)
    nbz = np.zeros(f2.value)
    nbd = np.zeros(f2.value, np.int32)
    rc = lib().ob_build_pattern(C.byref(prob), P(rp, _ip), P(ci, _ip), P(dg, _ip), P(nbp, _ip), P(nbc, _ip),
                                P(nbo, _ip), P(nbt), P(nba), P(nbz), P(nbd, _ip))
    assert rc == 0
    return dict(row_ptr=rp, col_idx=ci, diag_off=dg, nb_ptr=nbp, nb_cell=nbc, nb_off=nbo, nb_trans=nbt,
                nb_area=nba, nb_dzg=nbz, nb_dir=nbd)


def storage(prob, pv, mean):
    N = prob.nx * prob.ny * prob.nz
    st = np.zeros(3 * N)
    lib().ob_storage(C.byref(prob), P(pv), P(mean, _u8p), P(st))
    return st


def _dif(d):
    return C.byref(d) if d is not None else None


def linearize(prob, pv, mean, st1, dt, nthreads=1, diffusion=None, dispersion=None, nv=None):
    """(J, r); with dispersion: (J, r, nv_out), nv the cells' oil velocities of
    the previous linearization (zeros if None)."""
    N = prob.nx * prob.ny * prob.nz
    nz = C.c_int64()
    f2 = C.c_int64()
    assert lib().ob_pattern_size(C.byref(prob), C.byref(nz), C.byref(f2)) == 0
    r = np.zeros(3 * rows(prob))
    J = np.zeros(9 * nz.value)
    nv_in = nv_out = None
    if dispersion is not None:
        nv_in = np.zeros(N) if nv is None else np.ascontiguousarray(nv, dtype=np.float64)
        nv_out = np.zeros(N)
    rc = lib().ob_linearize_ext(C.byref(prob), _dif(diffusion), _dif(dispersion), P(nv_in), P(nv_out), P(pv),
                                P(mean, _u8p), P(st1), dt, P(r), P(J), nthreads)
    assert rc == 0
    return (J, r) if dispersion is None else (J, r, nv_out)


def linearize_domain(prob, pv, mean, st1, dt, cells, J, r):
    """linearizeDomain(domain) on cells over the previous (J, r); returns the
    new (J, r)"""
    J = np.array(J, dtype=np.float64, copy=True)
    r = np.array(r, dtype=np.float64, copy=True)
    c = np.ascontiguousarray(cells, dtype=np.int32).ravel()
    rc = lib().ob_linearize_domain(C.byref(prob), P(pv), P(mean, _u8p), P(st1), dt, c.size, P(c, _ip), P(r), P(J))
    assert rc == 0
    return J, r


def boundary_flows(prob, pv, mean):
    """(cell, face, flows [n][3]) of the boundary faces, cell and face order"""
    n = lib().ob_boundary_flows(C.byref(prob), P(pv), P(mean, _u8p), None, None, None)
    assert n >= 0
    cell = np.zeros(max(n, 1), np.int32)
    face = np.zeros(max(n, 1), np.int32)
    fl = np.zeros((max(n, 1), 3))
    assert lib().ob_boundary_flows(C.byref(prob), P(pv), P(mean, _u8p), P(cell, _ip), P(face, _ip), P(fl)) == n
    return cell[:n], face[:n], fl[:n]


def linearize_times(prob, pv, mean, st1, dt, nthreads, reps=5):
    """wall times of reps linearizations (the problem built once)"""
    t = np.zeros(reps)
    assert lib().ob_linearize_timed(C.byref(prob), P(pv), P(mean, _u8p), P(st1), dt, nthreads, reps, P(t)) == 0
    return t


def source_blocks(prob, pv, mean):
    """(cells, residual [n][3], diag [n][9]) of the problem's source terms."""
    n = prob.n_src
    cells = np.zeros(max(n, 1), np.int32)
    r = np.zeros(3 * max(n, 1))
    d = np.zeros(9 * max(n, 1))
    got = lib().ob_source_blocks(C.byref(prob), P(pv), P(mean, _u8p), P(cells, _ip), P(r), P(d))
    assert got == n
    return cells[:n], r[:3 * n], d[:9 * n]


def flows(prob, pv, mean, diffusion=None, dispersion=None, nv=None):
    """(flows, flores), [neighbour entries][3] in the pattern's order."""
    nz = C.c_int64()
    f2 = C.c_int64()
    lib().ob_pattern_size(C.byref(prob), C.byref(nz), C.byref(f2))
    N = prob.nx * prob.ny * prob.nz
    nf = nz.value - N
    fl = np.zeros((nf, 3))
    fr = np.zeros((nf, 3))
    nv = None if nv is None else np.ascontiguousarray(nv, dtype=np.float64)
    assert lib().ob_flows_ext(C.byref(prob), _dif(diffusion), _dif(dispersion), P(nv), P(pv), P(mean, _u8p), P(fl),
                              P(fr)) == 0
    return fl, fr


def intensive(prob, pv3, meaning, cell=0):
    out = np.zeros(13 * 4)
    pv3 = np.ascontiguousarray(pv3, dtype=np.float64)
    lib().ob_intensive(C.byref(prob), P(pv3), int(meaning), int(cell), P(out))
    names = ["p", "rho_w", "rho_o", "rho_g", "lam_w", "lam_o", "lam_g", "invB_w", "invB_o", "invB_g", "Rs",
             "phi", "So"]
    o = out.reshape(13, 4)
    return {n: o[k] for k, n in enumerate(names)}


def ilu0(n, bs, rp, ci, vals):
    lu = np.array(vals, dtype=np.float64, copy=True)
    rc = lib().ob_ilu0_factor(n, bs, P(rp, _ip), P(ci, _ip), P(lu))
    return rc, lu


def ilu0_apply(n, bs, rp, ci, lu, d, relax=1.0):
    v = np.zeros_like(d)
    lib().ob_ilu0_apply(n, bs, P(rp, _ip), P(ci, _ip), P(lu), P(d), P(v), relax)
    return v


def spmv(n, bs, rp, ci, vals, x):
    y = np.zeros(n * bs)
    lib().ob_spmv(n, bs, P(rp, _ip), P(ci, _ip), P(vals), P(x), P(y))
    return y


def colored_ilu_apply(dims, rp, ci, vals, d, relax=1.0):
    """Block ILU(0) in the two-colour tile ordering: factor, then M^-1 d."""
    nx, ny, nz = dims
    v = np.zeros(3 * nx * ny * nz)
    rc = lib().ob_colored_ilu(nx, ny, nz, P(rp, _ip), P(ci, _ip), P(np.ascontiguousarray(vals, dtype=np.float64)),
                              P(np.ascontiguousarray(d, dtype=np.float64)), P(v), relax)
    return rc, v


def bicgstab_grid(dims, rp, ci, vals, b, prm):
    nx, ny, nz = dims
    x = np.zeros(3 * nx * ny * nz)
    its = C.c_int32()
    rc = lib().ob_bicgstab_grid(nx, ny, nz, P(rp, _ip), P(ci, _ip), P(np.ascontiguousarray(vals, dtype=np.float64)),
                                P(np.ascontiguousarray(b, dtype=np.float64)), P(x), C.byref(prm), C.byref(its))
    return rc, x, its.value


def amg_params(**kw):
    """ob_params of the aggregation AMG (precond = 2); keywords as ob_params,
    amg_agg as a 3-tuple"""
    return params(precond=2, **kw)


def amg_dims(dims, prm):
    out = np.zeros((16, 3), dtype=np.int32)
    n = lib().ob_amg_dims(*dims, C.byref(prm), P(out, _ip), 16)
    return out[:max(n, 0)]


def amg_level_matrix(dims, rp, ci, vals, prm, level):
    """level `level`'s matrix of the AMG hierarchy, structured layout [7][N_l][9]"""
    d = amg_dims(dims, prm)
    A7 = np.zeros(63 * int(np.prod(d[level])))
    rc = lib().ob_amg_level_matrix(*dims, P(rp, _ip), P(ci, _ip), P(np.ascontiguousarray(vals, dtype=np.float64)),
                                   C.byref(prm), level, P(A7))
    return rc, A7


def amg_apply(dims, rp, ci, vals, prm, d):
    """one V-cycle of the AMG built from the matrix: v = M^-1 d"""
    v = np.zeros(3 * int(np.prod(dims)))
    rc = lib().ob_amg_apply(*dims, P(rp, _ip), P(ci, _ip), P(np.ascontiguousarray(vals, dtype=np.float64)),
                            C.byref(prm), P(np.ascontiguousarray(d, dtype=np.float64)), P(v))
    return rc, v


def gs_sweep(dims, A7, b, x0, forward=True, zero=False):
    """one block Gauss-Seidel sweep (two-colour tile order) of a structured-layout matrix"""
    x = np.array(x0, dtype=np.float64, copy=True)
    rc = lib().ob_gs_sweep(*dims, P(np.ascontiguousarray(A7, dtype=np.float64)),
                           P(np.ascontiguousarray(b, dtype=np.float64)), P(x), int(forward), int(zero))
    return rc, x


def dense_inverse(M):
    A = np.array(M, dtype=np.float64, copy=True, order="C")
    rc = lib().ob_dense_inverse(A.shape[0], P(A))
    return rc, A


def dia_from_bcrs(dims, rp, ci, vals):
    """the 7-point BCRS matrix in the structured layout [7][N][9] (absent blocks zero)"""
    nx, ny, nz = dims
    N = nx * ny * nz
    A7 = np.zeros((7, N, 9))
    vals = np.asarray(vals).reshape(-1, 9)
    off = (-nx * ny, -nx, -1, 0, 1, nx, nx * ny)
    for I in range(N):
        i, j, k = I % nx, (I // nx) % ny, I // (nx * ny)
        row = {int(ci[q]): q for q in range(rp[I], rp[I + 1])}
        for s, present in enumerate((k > 0, j > 0, i > 0, True, i + 1 < nx, j + 1 < ny, k + 1 < nz)):
            if present and I + off[s] in row:
                A7[s, I] = vals[row[I + off[s]]]
    return A7.reshape(-1)


def jacobi_apply(n, bs, rp, ci, vals, d, v0, order, relax=1.0):
    """`order` block-Jacobi sweeps from v0 (the preconditioner of bicgstab with precond=1)."""
    v = np.array(v0, dtype=np.float64, copy=True)
    rc = lib().ob_jacobi_apply(n, bs, P(rp, _ip), P(ci, _ip), P(np.ascontiguousarray(vals, dtype=np.float64)),
                               P(np.ascontiguousarray(d, dtype=np.float64)), P(v), order, relax)
    return rc, v


def dot(x, y):
    """The restated BiCGStab's scalar product: exact sum of x[i]*y[i], rounded once."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().ob_dot(x.size, P(x), P(y))


def bicgstab(n, bs, rp, ci, vals, b, prm):
    x = np.zeros(n * bs)
    its = C.c_int32()
    rc = lib().ob_bicgstab(n, bs, P(rp, _ip), P(ci, _ip), P(vals), P(b), P(x), C.byref(prm), C.byref(its))
    return rc, x, its.value


def update_primary(prob, cur, cur_mean, delta, was_switched, prm):
    N = prob.nx * prob.ny * prob.nz
    nxt = np.zeros(3 * N)
    nm = np.zeros(N, np.uint8)
    ws = np.array(was_switched, dtype=np.uint8, copy=True)
    ns = C.c_int32()
    rc = lib().ob_update_primary(C.byref(prob), P(cur), P(cur_mean, _u8p), P(delta), P(nxt), P(nm, _u8p),
                                 P(ws, _u8p), C.byref(prm), C.byref(ns))
    return rc, nxt, nm, ws, ns.value


def simulate(prob, pv0, mean0, n_steps, dt0, prm, nthreads=1, log_cap=4096, keep_history=True, diffusion=None,
             dispersion=None):
    N = prob.nx * prob.ny * prob.nz
    pv = np.array(pv0, dtype=np.float64, copy=True)
    mean = np.array(mean0, dtype=np.uint8, copy=True)
    hist = np.zeros((n_steps, 3 * N)) if keep_history else None
    dts = np.zeros(n_steps)
    log = (ObIterLog * log_cap)()
    nlog = C.c_int32()
    timers = np.zeros(4)
    rc = lib().ob_simulate_ext(C.byref(prob), _dif(diffusion), _dif(dispersion), P(pv), P(mean, _u8p), n_steps, dt0,
                               C.byref(prm), P(hist), P(dts), log, log_cap, C.byref(nlog), P(timers), nthreads)
    recs = [(l.step, l.attempt, l.iter, l.lin_iters, l.num_switched, l.lin_converged, l.error, l.dt)
            for l in log[:min(nlog.value, log_cap)]]
    return rc, pv, mean, hist, dts, recs, timers


OB_AUX_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, _dp, _u8p, _dp, _dp, _dp, C.c_int)


def simulate_aux(prob, pv0, mean0, aux0, n_steps, dt0, prm, module, nthreads=1, log_cap=4096):
    """simulate with auxiliary DOFs: module(pv, mean, aux, newton_iteration)
    returns the terms its linearize adds, (brow, bcol, blocks, rrow, rvals),
    or None when the linearization fails.  Returns rc, pv, mean, aux, hist,
    dts, log."""
    N = prob.nx * prob.ny * prob.nz
    na = 3 * prob.n_aux
    pv = np.array(pv0, dtype=np.float64, copy=True)
    mean = np.array(mean0, dtype=np.uint8, copy=True)
    aux = np.array(aux0, dtype=np.float64, copy=True)
    hist = np.zeros((n_steps, 3 * N))
    dts = np.zeros(n_steps)
    log = (ObIterLog * log_cap)()
    nlog = C.c_int32()

    def fn(user, pvp, mp, ap, jp, rp, it):
        terms = module(np.ctypeslib.as_array(pvp, (3 * N,)).copy(), np.ctypeslib.as_array(mp, (N,)).copy(),
                       np.ctypeslib.as_array(ap, (na,)).copy(), it)
        if terms is None:
            return 1
        brow, bcol, blocks, rrow, rvals = (np.ascontiguousarray(a) for a in terms)
        br = brow.astype(np.int32).ravel()
        bc = bcol.astype(np.int32).ravel()
        bl = blocks.astype(np.float64).ravel()
        rr = rrow.astype(np.int32).ravel()
        rv = rvals.astype(np.float64).ravel()
        return 0 if lib().ob_add_auxiliary(C.byref(prob), br.size, P(br, _ip), P(bc, _ip), P(bl), rr.size,
                                           P(rr, _ip), P(rv), jp, rp) == 0 else 1

    cb = OB_AUX_FN(fn)
    rc = lib().ob_simulate_aux(C.byref(prob), P(pv), P(mean, _u8p), P(aux), cb, None, n_steps, dt0, C.byref(prm),
                               P(hist), P(dts), log, log_cap, C.byref(nlog), nthreads)
    recs = [(l.step, l.attempt, l.iter, l.lin_iters, l.num_switched, l.lin_converged, l.error, l.dt)
            for l in log[:min(nlog.value, log_cap)]]
    return rc, pv, mean, aux, hist, dts, recs


# ---------------------------------------------------------------------------
# lens problem (two-phase immiscible element path)

def lens_problem(**kw):
    p = OlProblem()
    lib().ol_default_problem(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def lens_sizes(p):
    n, nnzb = C.c_int64(), C.c_int64()
    if lib().ol_num_cells(C.byref(p), C.byref(n), C.byref(nnzb)) != 0:
        raise RuntimeError("ol_num_cells failed")
    return n.value, nnzb.value


def lens_pattern(p):
    n, nnzb = lens_sizes(p)
    rp = np.zeros(n + 1, np.int32)
    ci = np.zeros(nnzb, np.int32)
    dg = np.zeros(n, np.int32)
    if lib().ol_pattern(C.byref(p), P(rp, _ip), P(ci, _ip), P(dg, _ip)) != 0:
        raise RuntimeError("ol_pattern failed")
    return rp, ci, dg


def lens_initial(p):
    n, _ = lens_sizes(p)
    pv = np.zeros(2 * n)
    lib().ol_initial(C.byref(p), P(pv))
    return pv


def lens_storage(p, pv):
    n, _ = lens_sizes(p)
    st = np.zeros(2 * n)
    pv = np.ascontiguousarray(pv, dtype=np.float64)
    if lib().ol_storage(C.byref(p), P(pv), P(st)) != 0:
        raise RuntimeError("ol_storage failed")
    return st


def lens_linearize(p, pv, st1, dt):
    n, nnzb = lens_sizes(p)
    res = np.zeros(2 * n)
    jac = np.zeros(4 * nnzb)
    pv = np.ascontiguousarray(pv, dtype=np.float64)
    st1 = np.ascontiguousarray(st1, dtype=np.float64)
    if lib().ol_linearize(C.byref(p), P(pv), P(st1), float(dt), P(res), P(jac)) != 0:
        raise RuntimeError("ol_linearize failed")
    return res, jac


def lens_material(p, lens, sw):
    out = np.zeros(6)
    lib().ol_material(C.byref(p), int(lens), float(sw), P(out))
    return dict(pc=out[0], dpc=out[1], krw=out[2], dkrw=out[3], krn=out[4], dkrn=out[5])


def lens_simulate(p, pv, t_end, dt0, prm, max_steps=10000, log_cap=100000):
    pv = np.array(pv, dtype=np.float64)
    th = np.zeros(max_steps)
    dth = np.zeros(max_steps)
    ns = C.c_int32()
    log = (ObIterLog * log_cap)()
    nlog = C.c_int32()
    rc = lib().ol_simulate(C.byref(p), P(pv), float(t_end), float(dt0), C.byref(prm), int(max_steps), P(th),
                           P(dth), C.byref(ns), log, log_cap, C.byref(nlog))
    recs = [(l.step, l.attempt, l.iter, l.lin_iters, l.num_switched, l.lin_converged, l.error, l.dt)
            for l in log[:min(nlog.value, log_cap)]]
    k = ns.value
    return rc, pv, th[:k], dth[:k], recs
