"""CPU checks of the C ABI library: it loads without a GPU and exports every
function declared in include/omodels.h with the binding's signature."""
import os
import re

import omodels_amd as om

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "omodels.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(omodels_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    lib = om.lib()
    names = declared_functions()
    assert len(names) > 20
    for n in sorted(names):
        assert hasattr(lib, n), n
    # the ctypes binding covers exactly the declared entry points
    assert names == set(om.SIGNATURES)


def test_host_only_entry_points_without_gpu():
    p = om.default_params()
    assert p.newton_max_iter == 20 and p.lin_tol == 1e-3 and p.dp_max_rel == 0.3
    oil, no, gas, ng, consts = om.default_fluid()
    assert no == 9 and ng == 10
    c = om.BlackOilCase(om.make_spec(4, 3, 2, seed=5))
    assert c.N == 24 and len(c.src_cell) == 4
