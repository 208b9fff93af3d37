"""The scalar product without allocation (omodels_scalar_product_noalloc):
device inputs from torch, the caller's stream, the accumulator in the
library's static device array.  Checked bit for bit against the allocating
entry point (omodels_scalar_product) and against math.fsum, on two calls in
a row, the second shorter, and on an oversized request that it must reject."""
import ctypes as C
import math

import numpy as np
import pytest

import omodels_amd as om

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _device_dot(x, y):
    dx = torch.from_numpy(x).to("cuda:0")
    dy = torch.from_numpy(y).to("cuda:0")
    out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda:0")
    stream = torch.cuda.current_stream(torch.device("cuda:0"))
    rc = om.lib().omodels_scalar_product_noalloc(x.size, C.c_void_p(dx.data_ptr()), C.c_void_p(dy.data_ptr()),
                                                 C.c_void_p(out.data_ptr()), C.c_void_p(stream.cuda_stream))
    assert rc == 0, om.lib().omodels_last_error().decode()
    torch.cuda.synchronize()
    return float(out.cpu().numpy()[0])


def _vectors(m, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(m) * np.resize([1e5, 1e-3, 1e2], m)
    y = rng.standard_normal(m) * np.resize([1e-4, 1e3, 1e-9], m)
    x[0], x[-1], y[0], y[-1] = 1e200, -1e200, 1.0, 1.0
    return x, y


def test_noalloc_matches_the_allocating_entry_point_twice():
    for m, seed in ((1_500_000, 1), (1001, 2)):
        x, y = _vectors(m, seed)
        got = _device_dot(x, y)
        want = om.scalar_product(x, y)
        assert got == want == math.fsum((x * y).tolist()), (m, got, want)


def test_noalloc_rejects_oversized_requests():
    rc = om.lib().omodels_scalar_product_noalloc(1 << 28, None, None, None, None)
    assert rc < 0
    assert "2^28" in om.lib().omodels_last_error().decode()
