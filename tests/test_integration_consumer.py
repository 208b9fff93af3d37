"""The C++ adapters of include/omodels_opm.hpp as reference-side code uses
them: tests/integration/consumer.cpp (a stand-in Simulator with its Glue)
compiles against include/ and libomodels_amd.so with -Wall -Wextra -Werror
on the CPU; on the GPU its Newton iteration (linearize, BiCGStab, update,
relinearize) through the adapters gives the bits of the same calls made
through the Python binding, on a distorted grid with threshold pressures."""
import os
import subprocess

import numpy as np
import pytest

import omodels_amd as om
from cases import case, random_discretization

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "integration", "consumer.cpp")
BUILT = os.path.join(HERE, "integration", "omodels_consumer")
DT = 43200.0


def compile_consumer(out):
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"), SRC,
           "-L" + os.path.join(ROOT, "opm-models_amd"), "-lomodels_amd",
           "-Wl,-rpath," + os.path.join(ROOT, "opm-models_amd"), "-o", out]
    subprocess.run(cmd, check=True, capture_output=True, text=True)


def write_case(c, d, geo=None):
    s = c.spec
    with open(os.path.join(d, "dims.txt"), "w") as f:
        f.write("%d %d %d %d %d %d %r %r %r %r %r\n" % (s.nx, s.ny, s.nz, len(c.src_cell), c.n_oil, c.n_gas, s.dx,
                                                        s.dy, s.dz, c.prob.gravity, DT))
    arrays = dict(perm=c.perm, poro=c.poro, oil_tab=c.oil, gas_tab=c.gas, consts=c.consts, src_cell=c.src_cell,
                  src_type=c.src_type, src_a=c.src_a, src_b=c.src_b, pv0=c.pv0, mean0=c.mean0)
    if geo is not None:
        arrays.update(zip(("trans", "depth", "volume", "thpres"), geo))
    for k, a in arrays.items():
        np.ascontiguousarray(a).tofile(os.path.join(d, k + ".bin"))


def test_consumer_compiles_and_reports_library_errors(tmp_path):
    exe = str(tmp_path / "omodels_consumer")
    compile_consumer(exe)
    c = case(4, 3, 2, seed=3)
    write_case(c, str(tmp_path))
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    # without a GPU the model cannot be created: the adapter throws with the
    # library's message; with one the iteration runs
    assert r.returncode in (0, 3), r.stderr
    # the host-only entry points answered before any device work
    assert '"lin_tol": 0.001' in r.stdout and '"n_oil": 9' in r.stdout and '"n_gas": 10' in r.stdout, r.stdout
    if r.returncode == 3:
        assert "model_create" in r.stderr and omodels_error_text(r.stderr)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stderr


def omodels_error_text(stderr):
    return len(stderr.split("model_create:", 1)[-1].strip()) > 0


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64 if a.dtype == np.float64 else a.dtype)


@pytest.mark.gpu
@pytest.mark.parametrize("distorted", [False, True], ids=["grid", "distorted"])
def test_consumer_newton_iteration_matches_the_binding(tmp_path, distorted):
    assert os.path.exists(BUILT), "build the consumer first (__graft_entry__.build())"
    c = case(12, 10, 4, seed=21)
    geo = random_discretization(c, 21) if distorted else None
    d = str(tmp_path)
    write_case(c, d, geo)
    r = subprocess.run([BUILT, d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = {k: np.fromfile(os.path.join(d, k + ".out"), dtype=np.uint8 if k == "mean1" else np.float64)
           for k in ("jac0", "res0", "x", "pv1", "mean1", "jac1", "res1")}
    # the same calls through the Python binding
    m = om.Model(c)
    if geo is not None:
        m.update_discretization(*geo)
    m.set_solution(c.pv0, c.mean0, both_levels=False)
    J0, r0 = m.linearize(DT, iteration_index=0)
    s = om.Solver(m, om.default_params())
    s.prepare(J0, r0)
    rc, x, its = s.solve()
    m.newton_update(x, om.default_params())
    pv1, mean1 = m.get_solution()
    J1, r1 = m.linearize(DT, iteration_index=1)
    for k, a in (("jac0", J0), ("res0", r0), ("x", x), ("pv1", pv1), ("jac1", J1), ("res1", r1)):
        assert np.array_equal(_bits(out[k]), _bits(a)), k
    assert np.array_equal(out["mean1"], mean1)
    assert '"converged": %d' % (1 if rc == 0 else 0) in r.stdout
    assert '"subdomain_same": 1' in r.stdout          # linearizeDomain(domain) over every cell
    # updateFlowsInfo, then linearize() with the consumer's auxiliary module
    fl, fr = m.flows()
    assert np.array_equal(_bits(np.fromfile(os.path.join(d, "flows.out"))), _bits(fl.ravel()))
    # velocities: FLORES over the face's area (axis from the neighbour's offset)
    pat = m.pattern()
    sp = c.spec
    area = np.array([sp.dy * sp.dz, sp.dx * sp.dz, sp.dx * sp.dy])
    vel = np.zeros_like(fr)
    for I in range(c.N):
        for k in range(pat["nb_ptr"][I], pat["nb_ptr"][I + 1]):
            dd = abs(int(pat["nb_cell"][k]) - I)
            vel[k] = fr[k] / area[0 if dd == 1 else (1 if dd == sp.nx else 2)]
    assert np.array_equal(_bits(np.fromfile(os.path.join(d, "velocity.out"))), _bits(vel.ravel()))
    N = m.N
    m.set_auxiliary(1, [N, 0, N], [N, N, 0])
    m.set_auxiliary_solution([1.0e5, 0.0, 0.0])
    J2, r2 = m.linearize(DT, iteration_index=1)
    k, dp = 1.0e-6, 1.0e5 - pv1[1]
    blocks = np.array([[1, 0, 0, 0, 1, 0, 0, 0, 1], [0, -1, 0, 0, 0, 0, 0, 0, 0],
                       [0, 0, 0, 0, -k, 0, 0, 0, 0], [0, 0, 0, k, 0, 0, 0, 0, 0]], dtype=np.float64)
    m.add_auxiliary([N, N, 0, 0], [N, 0, 0, N], blocks, [N, 0], [[dp, 0.0, 0.0], [0.0, k * dp, 0.0]])
    J2, r2 = m.system()
    assert np.array_equal(_bits(np.fromfile(os.path.join(d, "jac2.out"))), _bits(J2))
    assert np.array_equal(_bits(np.fromfile(os.path.join(d, "res2.out"))), _bits(r2))
