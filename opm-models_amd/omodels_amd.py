"""ctypes binding of libomodels_amd.so (the C ABI declared in include/omodels.h).

Host arrays are numpy arrays; every call returns plain numpy data.  The
library must have been built (python opm-models_amd/build_lib.py); there is
no fallback implementation: a missing library raises at load time.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libomodels_amd.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


class BlackOilProblem(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("n_src", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("gravity", C.c_double),
        ("perm", _dp), ("poro", _dp),
        ("n_oil", C.c_int32), ("n_gas", C.c_int32),
        ("oil_tab", _dp), ("gas_tab", _dp),
        ("oil_comp", C.c_double), ("water_pref", C.c_double), ("water_comp", C.c_double),
        ("water_mu", C.c_double), ("rho_ref", C.c_double * 3),
        ("src_cell", _ip), ("src_type", _ip), ("src_a", _dp), ("src_b", _dp),
    ]


class Diffusion(C.Structure):
    """omodels_diffusion: molecular diffusion in the oil phase."""
    _fields_ = [("oil_in_oil", C.c_double), ("gas_in_oil", C.c_double), ("mw_oil", C.c_double),
                ("mw_gas", C.c_double), ("face_diffusivity", _dp)]


class Dispersion(C.Structure):
    """omodels_dispersion: dispersion in the oil phase."""
    _fields_ = [("face_dispersivity", _dp), ("cell_dispersion", _dp)]


class BlackOilSpec(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("field", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("perm_mean_md", C.c_double), ("perm_sigma", C.c_double), ("kz_ratio", C.c_double),
        ("chan_perm_md", C.c_double), ("back_perm_md", C.c_double), ("chan_fraction", C.c_double),
        ("corr_len", C.c_double),
        ("poro_min", C.c_double), ("poro_max", C.c_double),
        ("p_top", C.c_double), ("sw_init", C.c_double),
        ("wells", C.c_int32), ("pad", C.c_int32),
        ("inj_rate", C.c_double), ("prod_bhp", C.c_double), ("seed", C.c_uint64), ("rs_frac", C.c_double),
    ]


class Params(C.Structure):
    _fields_ = [
        ("newton_tol", C.c_double), ("newton_max_iter", C.c_int32), ("newton_target_iter", C.c_int32),
        ("newton_max_error", C.c_double), ("lin_tol", C.c_double), ("lin_abs_tol", C.c_double),
        ("lin_max_error", C.c_double), ("lin_max_iter", C.c_int32), ("precond", C.c_int32),
        ("prec_relax", C.c_double), ("prec_order", C.c_int32), ("max_dt_divisions", C.c_int32),
        ("dp_max_rel", C.c_double), ("ds_max", C.c_double), ("privar_osc_thresh", C.c_double),
        ("max_dt", C.c_double), ("min_dt", C.c_double), ("precision", C.c_int32), ("verbose", C.c_int32),
        ("ilu_schedule", C.c_int32), ("autotune", C.c_int32), ("graph", C.c_int32), ("spmv_in_sweep", C.c_int32),
        ("ilu_order", C.c_int32), ("amg_coarsen_target", C.c_int32), ("amg_pre_smooth", C.c_int32),
        ("amg_post_smooth", C.c_int32), ("amg_prolong_damp", C.c_double), ("amg_agg", C.c_int32 * 3),
    ]


class IterLog(C.Structure):
    _fields_ = [("step", C.c_int32), ("attempt", C.c_int32), ("iter", C.c_int32), ("lin_iters", C.c_int32),
                ("num_switched", C.c_int32), ("lin_converged", C.c_int32), ("error", C.c_double),
                ("dt", C.c_double)]


class RunStats(C.Structure):
    _fields_ = [("newton_iters", C.c_int32), ("linearizations", C.c_int32), ("linear_iters", C.c_int32),
                ("failed_attempts", C.c_int32), ("t_linearize", C.c_double), ("t_solve", C.c_double),
                ("t_update", C.c_double), ("t_total", C.c_double)]


class LensProblem(C.Structure):
    """omodels_lens_problem: the two-phase immiscible lens problem."""
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


AUX_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_int)

# every symbol declared in include/omodels.h, with its ctypes signature
SIGNATURES = {
    "omodels_blackoil_default_fluid": (C.c_int, [_dp, _ip, _dp, _ip, _dp]),
    "omodels_blackoil_generate": (C.c_int, [C.POINTER(BlackOilSpec), _dp, _dp, _dp, _u8p, C.c_int32, _ip, _ip,
                                            _ip, _dp, _dp]),
    "omodels_default_params": (None, [C.POINTER(Params)]),
    "omodels_last_error": (C.c_char_p, []),
    "omodels_model_create": (C.c_void_p, [C.POINTER(BlackOilProblem)]),
    "omodels_model_destroy": (None, [C.c_void_p]),
    "omodels_model_num_cells": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "omodels_model_set_solution": (C.c_int, [C.c_void_p, _dp, _u8p, C.c_int]),
    "omodels_model_get_solution": (C.c_int, [C.c_void_p, _dp, _u8p]),
    "omodels_model_set_storage1": (C.c_int, [C.c_void_p, _dp]),
    "omodels_model_get_storage1": (C.c_int, [C.c_void_p, _dp]),
    "omodels_linearizer_pattern": (C.c_int, [C.c_void_p, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp]),
    "omodels_linearizer_linearize_domain": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    "omodels_linearizer_update_discretization": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
    "omodels_model_set_rock": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int32, _dp]),
    "omodels_linearizer_linearize_subdomain": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_int32, _ip]),
    "omodels_linearizer_get_jacobian": (C.c_int, [C.c_void_p, _dp]),
    "omodels_linearizer_get_residual": (C.c_int, [C.c_void_p, _dp]),
    "omodels_linearizer_erase_matrix": (C.c_int, [C.c_void_p]),
    "omodels_linearizer_set_separate_sparse_sources": (C.c_int, [C.c_void_p, C.c_int]),
    "omodels_linearizer_source_blocks": (C.c_int, [C.c_void_p, _ip, _dp, _dp]),
    "omodels_linearizer_add_sparse_sources": (C.c_int, [C.c_void_p, C.c_int32, _ip, _dp, _dp]),
    "omodels_linearizer_flows": (C.c_int, [C.c_void_p, _dp, _dp]),
    "omodels_model_set_diffusion": (C.c_int, [C.c_void_p, C.POINTER(Diffusion)]),
    "omodels_model_set_dispersion": (C.c_int, [C.c_void_p, C.POINTER(Dispersion)]),
    "omodels_model_get_cell_velocity": (C.c_int, [C.c_void_p, _dp]),
    "omodels_output_field_count": (C.c_int, []),
    "omodels_output_field_name": (C.c_char_p, [C.c_int32]),
    "omodels_output_fields": (C.c_int, [C.c_void_p, _dp]),
    "omodels_vtk_write": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                    C.c_int32, C.POINTER(C.c_char_p), _dp]),
    "omodels_model_write_vtk": (C.c_int, [C.c_void_p, C.c_char_p]),
    "omodels_restart_write": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, _dp, _u8p, C.c_double,
                                        C.c_int32]),
    "omodels_restart_read": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, _dp, _u8p,
                                       C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    "omodels_model_write_restart": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double, C.c_int32]),
    "omodels_model_read_restart": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double),
                                             C.POINTER(C.c_int32)]),
    "omodels_solver_create": (C.c_void_p, [C.c_void_p, C.POINTER(Params)]),
    "omodels_solver_destroy": (None, [C.c_void_p]),
    "omodels_solver_prepare": (C.c_int, [C.c_void_p, _dp, _dp]),
    "omodels_solver_get_residual": (C.c_int, [C.c_void_p, _dp]),
    "omodels_solver_solve": (C.c_int, [C.c_void_p, _dp]),
    "omodels_solver_iterations": (C.c_int, [C.c_void_p]),
    "omodels_solver_get_ilu": (C.c_int, [C.c_void_p, _dp]),
    "omodels_scalar_product": (C.c_int, [C.c_int64, _dp, _dp, _dp]),
    "omodels_scalar_product_noalloc": (C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "omodels_solver_apply_precond": (C.c_int, [C.c_void_p, _dp, _dp]),
    "omodels_model_set_auxiliary": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _ip, _ip]),
    "omodels_model_set_auxiliary_solution": (C.c_int, [C.c_void_p, _dp, C.c_int]),
    "omodels_model_get_auxiliary_solution": (C.c_int, [C.c_void_p, _dp]),
    "omodels_linearizer_add_auxiliary": (C.c_int, [C.c_void_p, C.c_int32, _ip, _ip, _dp, C.c_int32, _ip, _dp]),
    "omodels_model_set_auxiliary_module": (C.c_int, [C.c_void_p, AUX_FN, C.c_void_p]),
    "omodels_linearizer_boundary_flows": (C.c_int, [C.c_void_p, _ip, _ip, _dp]),
    "omodels_model_set_boundary_conditions": (C.c_int, [C.c_void_p, C.c_int32, _ip, _ip, _ip, _dp, _dp, _u8p, _dp, _dp]),
    "omodels_model_set_directional_mobility": (C.c_int, [C.c_void_p, _dp]),
    "omodels_solver_amg_levels": (C.c_int, [C.c_void_p, _ip, C.c_int]),
    "omodels_solver_get_amg_level": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "omodels_newton_update": (C.c_int, [C.c_void_p, C.POINTER(Params), _dp, _ip]),
    "omodels_model_get_switched": (C.c_int, [C.c_void_p, _u8p]),
    "omodels_lens_default_problem": (None, [C.POINTER(LensProblem)]),
    "omodels_lens_create": (C.c_void_p, [C.POINTER(LensProblem)]),
    "omodels_lens_destroy": (None, [C.c_void_p]),
    "omodels_lens_num_cells": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "omodels_lens_pattern": (C.c_int, [C.c_void_p, _ip, _ip, _ip]),
    "omodels_lens_initial": (C.c_int, [C.c_void_p, _dp]),
    "omodels_lens_set_solution": (C.c_int, [C.c_void_p, _dp, C.c_int]),
    "omodels_lens_get_solution": (C.c_int, [C.c_void_p, _dp]),
    "omodels_lens_set_storage1": (C.c_int, [C.c_void_p, _dp]),
    "omodels_lens_linearize": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    "omodels_lens_get_jacobian": (C.c_int, [C.c_void_p, _dp]),
    "omodels_lens_get_residual": (C.c_int, [C.c_void_p, _dp]),
    "omodels_lens_simulate": (C.c_int, [C.c_void_p, C.POINTER(Params), C.c_double, C.c_double, C.c_int, _dp, _dp,
                                        _ip, C.POINTER(IterLog), C.c_int, _ip, C.POINTER(RunStats)]),
    "omodels_server_create": (C.c_void_p, [C.POINTER(Params), C.c_double, C.c_int]),
    "omodels_server_destroy": (None, [C.c_void_p]),
    "omodels_server_submit": (C.c_int, [C.c_void_p, C.POINTER(BlackOilProblem), _dp, _u8p, C.c_int, C.c_double,
                                        C.c_double]),
    "omodels_server_run": (C.c_int, [C.c_void_p]),
    "omodels_server_result": (C.c_int, [C.c_void_p, C.c_int, _dp, _u8p, _dp, _dp, C.POINTER(RunStats)]),
    "omodels_server_log": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(IterLog), C.c_int, _ip]),
    "omodels_server_summary": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                         C.POINTER(C.c_int32)]),
    "omodels_server_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "omodels_server_inject_ilu_wait_error": (C.c_int, [C.c_void_p]),
    "omodels_server_kernel_times": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_int64), C.c_int]),
    "omodels_model_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "omodels_model_kernel_times": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_int64), C.c_int]),
    "omodels_simulate": (C.c_int, [C.c_void_p, C.POINTER(Params), C.c_int, C.c_double, _dp, _dp,
                                   C.POINTER(IterLog), C.c_int, _ip, C.POINTER(RunStats)]),
}

_lib = None


def lib():
    """Load the library once (raises OSError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libomodels_amd.so is not built: run opm-models_amd/build_lib.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a, t=_dp):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def scalar_product(x, y):
    """The backend's scalar product on the device: the exact sum of x[i]*y[i]
    rounded once (OverlappingScalarProduct::dot, overlappingscalarproduct.hh:58-70)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(1)
    check(lib().omodels_scalar_product(x.size, ptr(x), ptr(y), ptr(out)), "scalar_product")
    return float(out[0])


def output_field_names():
    return [lib().omodels_output_field_name(i).decode() for i in range(lib().omodels_output_field_count())]


def restart_file_name(output_dir, sim_name, t):
    """<dir>/<name>_time=<t>_rank=0.ers, the reference's restartFileName_ (restart.hh:80-95)."""
    d = "" if output_dir in (".", "") else output_dir.rstrip("/") + "/"
    return "%s%s_time=%s_rank=0.ers" % (d, sim_name, "%g" % t)


def write_restart(path, dims, pv, meaning, time, time_step_idx=0):
    nx, ny, nz = dims
    pv = np.ascontiguousarray(pv, dtype=np.float64)
    meaning = np.ascontiguousarray(meaning, dtype=np.uint8)
    check(lib().omodels_restart_write(os.fsencode(path), nx, ny, nz, ptr(pv), ptr(meaning, _u8p), float(time),
                                      int(time_step_idx)), "restart_write")


def read_restart(path, dims):
    """(pv [3N], meaning [N], time, time_step_idx)"""
    nx, ny, nz = dims
    n = nx * ny * nz
    pv = np.zeros(3 * n)
    meaning = np.zeros(n, np.uint8)
    t = C.c_double()
    k = C.c_int32()
    check(lib().omodels_restart_read(os.fsencode(path), nx, ny, nz, ptr(pv), ptr(meaning, _u8p), C.byref(t),
                                     C.byref(k)), "restart_read")
    return pv, meaning, t.value, k.value


def write_vtk(path, dims, spacing, fields):
    """fields: {name: [N] array} written as cell data of a .vtu file."""
    nx, ny, nz = dims
    names = list(fields)
    arr = np.ascontiguousarray(np.stack([np.asarray(fields[k], dtype=np.float64) for k in names]))
    cnames = (C.c_char_p * len(names))(*[n.encode() for n in names])
    check(lib().omodels_vtk_write(os.fsencode(path), nx, ny, nz, *map(float, spacing), len(names), cnames, ptr(arr)),
          "vtk_write")


class OmodelsError(RuntimeError):
    pass


def check(rc, what):
    if rc < 0:
        raise OmodelsError("%s failed (%d): %s" % (what, rc, lib().omodels_last_error().decode()))
    return rc


def default_params(**kw):
    p = Params()
    lib().omodels_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "amg_agg":
            for q in range(3):
                p.amg_agg[q] = int(v[q])
        else:
            setattr(p, k, v)
    return p


def default_fluid():
    oil = np.zeros(4 * 12)
    gas = np.zeros(3 * 12)
    no = C.c_int32()
    ng = C.c_int32()
    consts = np.zeros(7)
    check(lib().omodels_blackoil_default_fluid(ptr(oil), C.byref(no), ptr(gas), C.byref(ng), ptr(consts)),
          "default_fluid")
    return oil[:4 * no.value].copy(), no.value, gas[:3 * ng.value].copy(), ng.value, consts


def make_spec(nx, ny, nz, dx=50.0, dy=50.0, dz=5.0, field=0, perm_mean_md=100.0, perm_sigma=1.0,
              kz_ratio=0.1, chan_perm_md=1000.0, back_perm_md=1.0, chan_fraction=0.3, corr_len=0.0,
              poro_min=0.15, poro_max=0.30, p_top=200e5, sw_init=0.2, wells=0, inj_rate=None,
              prod_bhp=150e5, seed=1234, rs_frac=1.0):
    s = BlackOilSpec()
    s.nx, s.ny, s.nz, s.field = nx, ny, nz, field
    s.dx, s.dy, s.dz = dx, dy, dz
    s.perm_mean_md, s.perm_sigma, s.kz_ratio = perm_mean_md, perm_sigma, kz_ratio
    s.chan_perm_md, s.back_perm_md, s.chan_fraction, s.corr_len = chan_perm_md, back_perm_md, chan_fraction, corr_len
    s.poro_min, s.poro_max, s.p_top, s.sw_init = poro_min, poro_max, p_top, sw_init
    s.wells = wells
    if inj_rate is None:
        # 0.01% of the pore volume of water per day
        inj_rate = 1e-4 * (nx * dx * ny * dy * nz * dz) * 0.225 * 1037.0 / 86400.0
    s.inj_rate, s.prod_bhp, s.seed = inj_rate, prod_bhp, seed
    s.rs_frac = rs_frac
    return s


class BlackOilCase:
    """Host arrays of one black-oil problem plus the ctypes descriptor."""

    def __init__(self, spec):
        self.spec = spec
        n = spec.nx * spec.ny * spec.nz
        self.N = n
        self.perm = np.zeros(3 * n)
        self.poro = np.zeros(n)
        self.pv0 = np.zeros(3 * n)
        self.mean0 = np.zeros(n, dtype=np.uint8)
        cap = 8 * spec.nz + 8
        self.src_cell = np.zeros(cap, dtype=np.int32)
        self.src_type = np.zeros(cap, dtype=np.int32)
        self.src_a = np.zeros(cap)
        self.src_b = np.zeros(cap)
        ns = C.c_int32()
        check(lib().omodels_blackoil_generate(C.byref(spec), ptr(self.perm), ptr(self.poro), ptr(self.pv0),
                                              ptr(self.mean0, _u8p), cap, C.byref(ns), ptr(self.src_cell, _ip),
                                              ptr(self.src_type, _ip), ptr(self.src_a), ptr(self.src_b)),
              "generate")
        k = ns.value
        self.src_cell, self.src_type = self.src_cell[:k].copy(), self.src_type[:k].copy()
        self.src_a, self.src_b = self.src_a[:k].copy(), self.src_b[:k].copy()
        self.oil, self.n_oil, self.gas, self.n_gas, self.consts = default_fluid()
        self.prob = self.descriptor(BlackOilProblem)

    def descriptor(self, cls):
        s = self.spec
        p = cls()
        p.nx, p.ny, p.nz, p.n_src = s.nx, s.ny, s.nz, len(self.src_cell)
        p.dx, p.dy, p.dz, p.gravity = s.dx, s.dy, s.dz, -9.81
        p.perm, p.poro = ptr(self.perm), ptr(self.poro)
        p.n_oil, p.n_gas = self.n_oil, self.n_gas
        p.oil_tab, p.gas_tab = ptr(self.oil), ptr(self.gas)
        c = self.consts
        p.oil_comp, p.water_pref, p.water_comp, p.water_mu = c[0], c[1], c[2], c[3]
        p.rho_ref[0], p.rho_ref[1], p.rho_ref[2] = c[4], c[5], c[6]
        p.src_cell, p.src_type = ptr(self.src_cell, _ip), ptr(self.src_type, _ip)
        p.src_a, p.src_b = ptr(self.src_a), ptr(self.src_b)
        return p


class Model:
    """Device-resident reservoir state with the linearizer interface."""

    def __init__(self, case):
        self.case = case
        h = lib().omodels_model_create(C.byref(case.prob))
        if not h:
            raise OmodelsError("model_create failed: " + lib().omodels_last_error().decode())
        self.h = C.c_void_p(h)
        n = C.c_int64()
        z = C.c_int64()
        lib().omodels_model_num_cells(self.h, C.byref(n), C.byref(z))
        self.N, self.nnzb = n.value, z.value
        self.grid_nnzb = self.nnzb
        self.n_aux = 0
        self._aux_fn = None

    @property
    def rows(self):
        """rows of the system: the cells and the auxiliary DOFs"""
        return self.N + self.n_aux

    def close(self):
        if self.h:
            lib().omodels_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_solution(self, pv, mean, both_levels=True):
        pv = np.ascontiguousarray(pv, dtype=np.float64)
        mean = np.ascontiguousarray(mean, dtype=np.uint8)
        check(lib().omodels_model_set_solution(self.h, ptr(pv), ptr(mean, _u8p), int(both_levels)), "set_solution")

    def get_solution(self):
        pv = np.zeros(3 * self.N)
        m = np.zeros(self.N, dtype=np.uint8)
        check(lib().omodels_model_get_solution(self.h, ptr(pv), ptr(m, _u8p)), "get_solution")
        return pv, m

    def set_storage1(self, st=None):
        check(lib().omodels_model_set_storage1(self.h, ptr(st) if st is not None else None), "set_storage1")

    def get_storage1(self):
        st = np.zeros(3 * self.N)
        check(lib().omodels_model_get_storage1(self.h, ptr(st)), "get_storage1")
        return st

    def pattern(self):
        N, z, R = self.N, self.nnzb, self.rows
        nf = self.grid_nnzb - N
        rp = np.zeros(R + 1, np.int32)
        ci = np.zeros(z, np.int32)
        dg = np.zeros(R, np.int32)
        nbp = np.zeros(N + 1, np.int32)
        nbc = np.zeros(nf, np.int32)
        nbo = np.zeros(nf, np.int32)
        nbt = np.zeros(nf)
        nbz = np.zeros(nf)
        check(lib().omodels_linearizer_pattern(self.h, ptr(rp, _ip), ptr(ci, _ip), ptr(dg, _ip), ptr(nbp, _ip),
                                               ptr(nbc, _ip), ptr(nbo, _ip), ptr(nbt), ptr(nbz)), "pattern")
        return dict(row_ptr=rp, col_idx=ci, diag_off=dg, nb_ptr=nbp, nb_cell=nbc, nb_off=nbo, nb_trans=nbt,
                    nb_dzg=nbz)

    def update_discretization(self, trans=None, depth=None, volume=None, thpres=None):
        """updateStoredTransmissibilities and the other per-neighbour inputs:
        trans / thpres [3][N] at the face's index (axis*N + I for face
        I|I+e_axis), depth / volume [N]; None for the grid's own."""
        N = self.N
        arrs = []
        for a, n in ((trans, 3 * N), (depth, N), (volume, N), (thpres, 3 * N)):
            if a is None:
                arrs.append(None)
                continue
            a = np.ascontiguousarray(a, dtype=np.float64).ravel()
            if a.size != n:
                raise ValueError("update_discretization: array of %d values, expected %d" % (a.size, n))
            arrs.append(a)
        self._geo = arrs        # the library copies them; kept for the caller's inspection
        check(lib().omodels_linearizer_update_discretization(
            self.h, *[ptr(a) if a is not None else None for a in arrs]), "update_discretization")

    def set_rock(self, comp=None, pref=None, table=None):
        """rock compressibility [N] (1/Pa), reference pressures [N] (Pa) and
        the multiplier table [3][n] (p_o, pore-volume multiplier,
        transmissibility multiplier); None for the problem's defaults."""
        N = self.N
        c = None if comp is None else np.ascontiguousarray(comp, dtype=np.float64).ravel()
        p = None if pref is None else np.ascontiguousarray(pref, dtype=np.float64).ravel()
        t = None if table is None else np.ascontiguousarray(table, dtype=np.float64).ravel()
        for a in (c, p):
            if a is not None and a.size != N:
                raise ValueError("set_rock: per-cell arrays of %d values" % N)
        if t is not None and t.size % 3:
            raise ValueError("set_rock: the table holds three rows")
        n = 0 if t is None else t.size // 3
        check(lib().omodels_model_set_rock(self.h, ptr(c), ptr(p), n, ptr(t)), "set_rock")

    def set_boundary(self, cell=None, face=None, type=None, rate=None, state=None, meaning=None, trans=None,
                     zface=None):
        """boundary conditions on boundary faces (see
        omodels_model_set_boundary_conditions); cell None removes them"""
        if cell is None:
            check(lib().omodels_model_set_boundary_conditions(self.h, 0, None, None, None, None, None, None, None,
                                                              None), "set_boundary")
            self._bc = None
            return
        keep = dict(cell=np.ascontiguousarray(cell, dtype=np.int32).ravel(),
                    face=np.ascontiguousarray(face, dtype=np.int32).ravel(),
                    type=np.ascontiguousarray(type, dtype=np.int32).ravel())
        for k, a, dt in (("rate", rate, np.float64), ("state", state, np.float64), ("meaning", meaning, np.uint8),
                         ("trans", trans, np.float64), ("zface", zface, np.float64)):
            keep[k] = None if a is None else np.ascontiguousarray(a, dtype=dt).ravel()
        n = keep["cell"].size
        check(lib().omodels_model_set_boundary_conditions(
            self.h, n, ptr(keep["cell"], _ip), ptr(keep["face"], _ip), ptr(keep["type"], _ip), ptr(keep["rate"]),
            ptr(keep["state"]), ptr(keep["meaning"], _u8p), ptr(keep["trans"]), ptr(keep["zface"])), "set_boundary")
        self._bc = keep

    def set_auxiliary(self, n_dofs=0, rows=None, cols=None):
        """auxiliary DOFs after the cells and the pattern entries (rows[k],
        cols[k]) their module adds; 0 for none"""
        r = np.ascontiguousarray([] if rows is None else rows, dtype=np.int32).ravel()
        c = np.ascontiguousarray([] if cols is None else cols, dtype=np.int32).ravel()
        check(lib().omodels_model_set_auxiliary(self.h, int(n_dofs), r.size, ptr(r, _ip), ptr(c, _ip)), "set_auxiliary")
        n = C.c_int64()
        z = C.c_int64()
        lib().omodels_model_num_cells(self.h, C.byref(n), C.byref(z))
        self.nnzb = z.value
        self.n_aux = int(n_dofs)

    def set_auxiliary_solution(self, x, both_levels=True):
        x = np.ascontiguousarray(x, dtype=np.float64)
        check(lib().omodels_model_set_auxiliary_solution(self.h, ptr(x), int(both_levels)), "set_auxiliary_solution")

    def get_auxiliary_solution(self):
        x = np.zeros(3 * self.n_aux)
        check(lib().omodels_model_get_auxiliary_solution(self.h, ptr(x)), "get_auxiliary_solution")
        return x

    def add_auxiliary(self, brow=(), bcol=(), blocks=None, rrow=(), rvals=None):
        br = np.ascontiguousarray(brow, dtype=np.int32).ravel()
        bc = np.ascontiguousarray(bcol, dtype=np.int32).ravel()
        bl = np.ascontiguousarray(np.zeros(0) if blocks is None else blocks, dtype=np.float64).ravel()
        rr = np.ascontiguousarray(rrow, dtype=np.int32).ravel()
        rv = np.ascontiguousarray(np.zeros(0) if rvals is None else rvals, dtype=np.float64).ravel()
        check(lib().omodels_linearizer_add_auxiliary(self.h, br.size, ptr(br, _ip), ptr(bc, _ip), ptr(bl), rr.size,
                                                     ptr(rr, _ip), ptr(rv)), "add_auxiliary")

    def set_auxiliary_module(self, fn=None):
        """fn(model, newton_iteration) -> 0 adds the module's terms after each
        linearization of simulate (nonzero: the linearization failed)"""
        if fn is None:
            self._aux_fn = None
            check(lib().omodels_model_set_auxiliary_module(self.h, AUX_FN(), None), "set_auxiliary_module")
            return

        def tramp(user, handle, it):
            try:
                return int(fn(self, it))
            except Exception:
                return 1
        self._aux_fn = AUX_FN(tramp)
        check(lib().omodels_model_set_auxiliary_module(self.h, self._aux_fn, None), "set_auxiliary_module")

    def boundary_flows(self):
        """(cell, face, flows [n][3]) of the boundary faces with a condition,
        in cell and face order"""
        n = check(lib().omodels_linearizer_boundary_flows(self.h, None, None, None), "boundary_flows")
        cell = np.zeros(n, np.int32)
        face = np.zeros(n, np.int32)
        fl = np.zeros((n, 3))
        if n:
            check(lib().omodels_linearizer_boundary_flows(self.h, ptr(cell, _ip), ptr(face, _ip), ptr(fl)),
                  "boundary_flows")
        return cell, face, fl

    def set_directional_mobility(self, mult=None):
        """directional relative permeability multipliers [3][3][N] (axis x, y,
        z; phase water, oil, gas; cell); None switches them off"""
        m = None if mult is None else np.ascontiguousarray(mult, dtype=np.float64).ravel()
        if m is not None and m.size != 9 * self.N:
            raise ValueError("set_directional_mobility: [3][3][N] multipliers")
        self._dirmob = m
        check(lib().omodels_model_set_directional_mobility(self.h, ptr(m)), "set_directional_mobility")

    def linearize(self, dt, iteration_index=1):
        check(lib().omodels_linearizer_linearize_domain(self.h, dt, iteration_index), "linearize_domain")
        J = np.zeros(9 * self.nnzb)
        r = np.zeros(3 * self.rows)
        check(lib().omodels_linearizer_get_jacobian(self.h, ptr(J)), "get_jacobian")
        check(lib().omodels_linearizer_get_residual(self.h, ptr(r)), "get_residual")
        return J, r

    def linearize_subdomain(self, dt, cells, iteration_index=1):
        """linearizeDomain(domain) on the given cells; returns the whole (J, r)"""
        c = np.ascontiguousarray(cells, dtype=np.int32).ravel()
        check(lib().omodels_linearizer_linearize_subdomain(self.h, dt, iteration_index, c.size, ptr(c, _ip)),
              "linearize_subdomain")
        return self.system()

    def system(self):
        """The linearizer's current Jacobian values and residual (no relinearization)."""
        J = np.zeros(9 * self.nnzb)
        r = np.zeros(3 * self.rows)
        check(lib().omodels_linearizer_get_jacobian(self.h, ptr(J)), "get_jacobian")
        check(lib().omodels_linearizer_get_residual(self.h, ptr(r)), "get_residual")
        return J, r

    def set_separate_sparse_sources(self, enable=True):
        check(lib().omodels_linearizer_set_separate_sparse_sources(self.h, int(enable)), "set_separate_sparse_sources")

    def source_blocks(self):
        """(cells, residual [n][3], diag [n][9]) of the problem's source terms at solution(0)."""
        n = self.case.prob.n_src
        cells = np.zeros(max(n, 1), np.int32)
        r = np.zeros(3 * max(n, 1))
        d = np.zeros(9 * max(n, 1))
        got = check(lib().omodels_linearizer_source_blocks(self.h, ptr(cells, _ip), ptr(r), ptr(d)), "source_blocks")
        return cells[:got], r[:3 * got], d[:9 * got]

    def add_sparse_sources(self, cells, residual, diag):
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        residual = np.ascontiguousarray(residual, dtype=np.float64)
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        check(lib().omodels_linearizer_add_sparse_sources(self.h, len(cells), ptr(cells, _ip), ptr(residual), ptr(diag)),
              "add_sparse_sources")

    def output_fields(self):
        """{name: [N] array} of the output quantities of solution(0)."""
        names = output_field_names()
        out = np.zeros((len(names), self.N))
        check(lib().omodels_output_fields(self.h, ptr(out)), "output_fields")
        return {n: out[i] for i, n in enumerate(names)}

    def write_vtk(self, path):
        check(lib().omodels_model_write_vtk(self.h, os.fsencode(path)), "write_vtk")

    def write_restart(self, path, time, time_step_idx=0):
        check(lib().omodels_model_write_restart(self.h, os.fsencode(path), float(time), int(time_step_idx)),
              "write_restart")

    def read_restart(self, path):
        """Loads solution(0) and solution(1); returns (time, time_step_idx)."""
        t = C.c_double()
        k = C.c_int32()
        check(lib().omodels_model_read_restart(self.h, os.fsencode(path), C.byref(t), C.byref(k)), "read_restart")
        return t.value, k.value

    def flows(self):
        """(flows, flores), each [nnzb - N][3] in the pattern's neighbour order, at solution(0)."""
        nf = self.grid_nnzb - self.N
        fl = np.zeros((nf, 3))
        fr = np.zeros((nf, 3))
        check(lib().omodels_linearizer_flows(self.h, ptr(fl), ptr(fr)), "flows")
        return fl, fr

    def set_diffusion(self, oil_in_oil=None, gas_in_oil=0.0, mw_oil=0.0, mw_gas=0.0, face_diffusivity=None):
        """Molecular diffusion in the oil phase for the following linearizations
        (coefficients in m^2/s; face_diffusivity [3][N] or None to derive it
        from the porosity); oil_in_oil=None switches it off."""
        if oil_in_oil is None:
            check(lib().omodels_model_set_diffusion(self.h, None), "set_diffusion")
            return
        d = Diffusion(float(oil_in_oil), float(gas_in_oil), float(mw_oil), float(mw_gas), None)
        if face_diffusivity is not None:
            fd = np.ascontiguousarray(face_diffusivity, dtype=np.float64)
            if fd.size != 3 * self.N:
                raise ValueError("face_diffusivity must hold 3*N values")
            d.face_diffusivity = ptr(fd)
        check(lib().omodels_model_set_diffusion(self.h, C.byref(d)), "set_diffusion")

    def set_dispersion(self, face_dispersivity=None, cell_dispersion=None, off=False):
        """Dispersion in the oil phase for the following linearizations (face
        dispersivities [3][N] or cell coefficients [N]); off=True switches it off."""
        if off:
            check(lib().omodels_model_set_dispersion(self.h, None), "set_dispersion")
            return
        d = Dispersion(None, None)
        keep = []
        for name, a, n in (("face_dispersivity", face_dispersivity, 3 * self.N),
                           ("cell_dispersion", cell_dispersion, self.N)):
            if a is not None:
                a = np.ascontiguousarray(a, dtype=np.float64)
                if a.size != n:
                    raise ValueError("%s must hold %d values" % (name, n))
                setattr(d, name, ptr(a))
                keep.append(a)
        check(lib().omodels_model_set_dispersion(self.h, C.byref(d)), "set_dispersion")

    def cell_velocity(self):
        """The cells' oil velocities the next linearization uses (dispersion on)."""
        v = np.zeros(self.N)
        check(lib().omodels_model_get_cell_velocity(self.h, ptr(v)), "get_cell_velocity")
        return v

    def set_profiling(self, enable=True):
        check(lib().omodels_model_set_profiling(self.h, int(enable)), "set_profiling")

    KERNEL_CLASSES = ("assemble", "spmv", "ilu_apply", "ilu_factor", "other")

    def kernel_times(self):
        ms = np.zeros(5)
        cnt = np.zeros(5, np.int64)
        check(lib().omodels_model_kernel_times(self.h, ptr(ms), cnt.ctypes.data_as(C.POINTER(C.c_int64)), 5),
              "kernel_times")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(self.KERNEL_CLASSES)}

    def newton_update(self, delta, params):
        delta = np.ascontiguousarray(delta, dtype=np.float64)
        ns = C.c_int32()
        rc = check(lib().omodels_newton_update(self.h, C.byref(params), ptr(delta), C.byref(ns)), "newton_update")
        sw = np.zeros(self.N, np.uint8)
        check(lib().omodels_model_get_switched(self.h, ptr(sw, _u8p)), "get_switched")
        return rc, ns.value, sw

    def simulate(self, params, n_steps, dt0, keep_history=True, log_cap=4096):
        hist = np.zeros((n_steps, 3 * self.N)) if keep_history else None
        dts = np.zeros(n_steps)
        log = (IterLog * log_cap)()
        nlog = C.c_int32()
        st = RunStats()
        rc = check(lib().omodels_simulate(self.h, C.byref(params), n_steps, dt0,
                                          ptr(hist) if hist is not None else None, ptr(dts), log, log_cap,
                                          C.byref(nlog), C.byref(st)), "simulate")
        recs = [(l.step, l.attempt, l.iter, l.lin_iters, l.num_switched, l.lin_converged, l.error, l.dt)
                for l in log[:min(nlog.value, log_cap)]]
        return rc, hist, dts, recs, st


class Solver:
    """BiCGStab backend bound to a model."""

    def __init__(self, model, params):
        self.model = model
        h = lib().omodels_solver_create(model.h, C.byref(params))
        if not h:
            raise OmodelsError("solver_create failed: " + lib().omodels_last_error().decode())
        self.h = C.c_void_p(h)

    def close(self):
        if self.h:
            lib().omodels_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self, values=None, residual=None):
        check(lib().omodels_solver_prepare(self.h, ptr(values), ptr(residual)), "solver_prepare")

    def solve(self):
        x = np.zeros(3 * self.model.rows)
        rc = check(lib().omodels_solver_solve(self.h, ptr(x)), "solver_solve")
        return rc, x, lib().omodels_solver_iterations(self.h)

    def ilu(self):
        lu = np.zeros(9 * self.model.nnzb)
        check(lib().omodels_solver_get_ilu(self.h, ptr(lu)), "get_ilu")
        return lu

    def apply_precond(self, d):
        d = np.ascontiguousarray(d, dtype=np.float64)
        v = np.zeros_like(d)
        check(lib().omodels_solver_apply_precond(self.h, ptr(d), ptr(v)), "apply_precond")
        return v

    def amg_dims(self):
        """level dimensions of the AMG hierarchy (precond = 2), [levels][3]"""
        dims = np.zeros((16, 3), dtype=np.int32)
        n = check(lib().omodels_solver_amg_levels(self.h, ptr(dims, _ip), 16), "amg_levels")
        return dims[:n]

    def amg_level(self, level):
        """level `level`'s matrix in the structured layout [7][N_l][9]"""
        d = self.amg_dims()[level]
        a = np.zeros(63 * int(np.prod(d)))
        check(lib().omodels_solver_get_amg_level(self.h, level, ptr(a)), "get_amg_level")
        return a


class Server:
    """Ensemble engine behind a request loop (continuous batching)."""

    def __init__(self, params, hbm_budget_bytes=64e9, max_batch=256):
        h = lib().omodels_server_create(C.byref(params), float(hbm_budget_bytes), int(max_batch))
        if not h:
            raise OmodelsError("server_create failed: " + lib().omodels_last_error().decode())
        self.h = C.c_void_p(h)
        self.cases = {}

    def close(self):
        if self.h:
            lib().omodels_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, case, n_steps, dt0=86400.0, arrival=0.0, pv=None, mean=None):
        pv = case.pv0 if pv is None else np.ascontiguousarray(pv, dtype=np.float64)
        mean = case.mean0 if mean is None else np.ascontiguousarray(mean, dtype=np.uint8)
        rid = check(lib().omodels_server_submit(self.h, C.byref(case.prob), ptr(pv), ptr(mean, _u8p), int(n_steps),
                                                float(dt0), float(arrival)), "server_submit")
        self.cases[rid] = (case, n_steps)
        return rid

    def run(self):
        check(lib().omodels_server_run(self.h), "server_run")

    def result(self, rid):
        case, n_steps = self.cases[rid]
        pv = np.zeros(3 * case.N)
        mean = np.zeros(case.N, np.uint8)
        times = np.zeros(3)
        dts = np.zeros(n_steps)
        st = RunStats()
        status = check(lib().omodels_server_result(self.h, rid, ptr(pv), ptr(mean, _u8p), ptr(times), ptr(dts),
                                                   C.byref(st)), "server_result")
        return dict(status=status, pv=pv, mean=mean, arrival=times[0], admit=times[1], finish=times[2], dts=dts,
                    stats=st)

    def log(self, rid, cap=4096):
        buf = (IterLog * cap)()
        n = C.c_int32()
        check(lib().omodels_server_log(self.h, rid, buf, cap, C.byref(n)), "server_log")
        return [(l.step, l.attempt, l.iter, l.lin_iters, l.num_switched, l.lin_converged, l.error, l.dt)
                for l in buf[:min(n.value, cap)]]

    def summary(self):
        w = C.c_double()
        bs = C.c_int64()
        mr = C.c_int32()
        check(lib().omodels_server_summary(self.h, C.byref(w), C.byref(bs), C.byref(mr)), "server_summary")
        return dict(wall=w.value, batch_rounds=bs.value, max_running=mr.value)

    def set_profiling(self, enable=True):
        check(lib().omodels_server_set_profiling(self.h, int(enable)), "server_set_profiling")

    def kernel_times(self):
        ms = np.zeros(5)
        cnt = np.zeros(5, np.int64)
        check(lib().omodels_server_kernel_times(self.h, ptr(ms), cnt.ctypes.data_as(C.POINTER(C.c_int64)), 5),
              "server_kernel_times")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(Model.KERNEL_CLASSES)}


# ---------------------------------------------------------------------------
# two-phase immiscible element path (the lens problem)

def lens_problem(**kw):
    p = LensProblem()
    lib().omodels_lens_default_problem(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class LensModel:
    """Device-resident lens problem with the linearizer interface and the Newton / time loop."""

    def __init__(self, prob=None):
        self.prob = prob if prob is not None else lens_problem()
        h = lib().omodels_lens_create(C.byref(self.prob))
        if not h:
            raise OmodelsError("lens_create failed: " + lib().omodels_last_error().decode())
        self.h = C.c_void_p(h)
        n = C.c_int64()
        z = C.c_int64()
        lib().omodels_lens_num_cells(self.h, C.byref(n), C.byref(z))
        self.N, self.nnzb = n.value, z.value

    def close(self):
        if self.h:
            lib().omodels_lens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pattern(self):
        rp = np.zeros(self.N + 1, np.int32)
        ci = np.zeros(self.nnzb, np.int32)
        dg = np.zeros(self.N, np.int32)
        check(lib().omodels_lens_pattern(self.h, ptr(rp, _ip), ptr(ci, _ip), ptr(dg, _ip)), "lens_pattern")
        return rp, ci, dg

    def initial(self):
        pv = np.zeros(2 * self.N)
        check(lib().omodels_lens_initial(self.h, ptr(pv)), "lens_initial")
        return pv

    def set_solution(self, pv, both_levels=True):
        pv = np.ascontiguousarray(pv, dtype=np.float64)
        check(lib().omodels_lens_set_solution(self.h, ptr(pv), int(both_levels)), "lens_set_solution")

    def get_solution(self):
        pv = np.zeros(2 * self.N)
        check(lib().omodels_lens_get_solution(self.h, ptr(pv)), "lens_get_solution")
        return pv

    def set_storage1(self, st):
        st = np.ascontiguousarray(st, dtype=np.float64)
        check(lib().omodels_lens_set_storage1(self.h, ptr(st)), "lens_set_storage1")

    def linearize(self, dt, iteration_index=1):
        check(lib().omodels_lens_linearize(self.h, float(dt), int(iteration_index)), "lens_linearize")
        r = np.zeros(2 * self.N)
        J = np.zeros(4 * self.nnzb)
        check(lib().omodels_lens_get_residual(self.h, ptr(r)), "lens_get_residual")
        check(lib().omodels_lens_get_jacobian(self.h, ptr(J)), "lens_get_jacobian")
        return r, J

    def simulate(self, params, t_end, dt0, max_steps=10000, log_cap=100000):
        th = np.zeros(max_steps)
        dth = np.zeros(max_steps)
        ns = C.c_int32()
        log = (IterLog * log_cap)()
        nlog = C.c_int32()
        st = RunStats()
        rc = check(lib().omodels_lens_simulate(self.h, C.byref(params), float(t_end), float(dt0), int(max_steps),
                                               ptr(th), ptr(dth), C.byref(ns), log, log_cap, C.byref(nlog),
                                               C.byref(st)), "lens_simulate")
        recs = [(l.step, l.attempt, l.iter, l.lin_iters, l.num_switched, l.lin_converged, l.error, l.dt)
                for l in log[:min(nlog.value, log_cap)]]
        k = ns.value
        return rc, th[:k], dth[:k], recs, st
