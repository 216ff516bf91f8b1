"""The allocation-free parallel forward projector (GPU test).

ccpi_pb_f2D_noalloc keeps the geometry's tables in a fixed device array
instead of a cached plan.  Two calls in a row through the C ABI, the second
on a smaller geometry (so the array holds the first call's larger tables
when the second call runs), must each equal ccpi_pb_f2D on the same input
bit for bit: both run the same kernel on the same tables."""

import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from ccpi.reconstruction import _lib
from geometries import centred, ellipsoid_phantom

pytestmark = pytest.mark.gpu


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _case(n, nz, na, nh, shift, seed):
    ang = np.linspace(0, np.pi, na, endpoint=False) + 0.01
    h = centred(nh, 1.0, shift)
    v = centred(nz, 1.0)
    o = np.array([-n / 2.0, -n / 2.0, -nz / 2.0])
    s = np.array([1.0, 1.0, 1.0])
    x = ellipsoid_phantom((n, n, nz), s, o, 5, seed=seed)
    return ang, h, v, o, s, x


def _run(fn, ang, h, v, o, s, x, stream):
    n, _, nz = x.shape
    vol = torch.from_numpy(x).cuda()
    out = torch.zeros((ang.size, h.size, v.size), dtype=torch.float32, device="cuda")
    _lib.check(fn(_dp(h), _dp(v), _dp(ang), ang.size, h.size, v.size, _dp(o), _dp(s), n, n, nz,
                  C.c_void_p(out.data_ptr()), C.c_void_p(vol.data_ptr()), stream))
    return out


def test_noalloc_fp_matches_plan_fp_on_two_calls():
    L = _lib.lib()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    big = _case(96, 12, 180, 137, 0.3, seed=1)
    small = _case(40, 5, 45, 57, -0.2, seed=2)
    got = [_run(L.ccpi_pb_f2D_noalloc, *big, stream), _run(L.ccpi_pb_f2D_noalloc, *small, stream)]
    want = [_run(L.ccpi_pb_f2D, *big, stream), _run(L.ccpi_pb_f2D, *small, stream)]
    torch.cuda.synchronize()
    for g, w in zip(got, want):
        assert float(w.abs().max()) > 0
        assert torch.equal(g, w)
