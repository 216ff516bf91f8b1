"""CPU checks of the C ABI: the library loads, exports every function
include/ccpi_hip.h declares, and rejects invalid geometry with status codes
before touching a device (no compute calls here)."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from ccpi.reconstruction import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "ccpi_hip.h")).read()
    return sorted(set(re.findall(r"^\w[\w \*]*?\b(ccpi_\w+)\s*\(", text, re.M)))


def test_header_and_binding_agree():
    assert sorted(_lib.EXPORTS) == _declared()


def test_library_exports_every_symbol():
    L = _lib.lib()
    for name in _declared():
        assert hasattr(L, name), name


def test_utilities_without_device():
    L = _lib.lib()
    assert L.ccpi_version() >= 100
    assert L.ccpi_error_string(-2).decode().startswith("geometry")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@pytest.mark.parametrize("case", ["one_column", "decreasing_h", "zero_voxel"])
def test_parallel_geometry_rejected(case):
    L = _lib.lib()
    h = np.array([-1.0, 0.0, 1.0])
    v = np.array([0.5])
    ang = np.array([0.0, 0.5])
    o = np.array([-1.0, -1.0, 0.0])
    s = np.array([1.0, 1.0, 1.0])
    if case == "one_column":
        h = h[:1]
    elif case == "decreasing_h":
        h = h[::-1].copy()
    else:
        s = np.array([1.0, 0.0, 1.0])
    dummy = C.c_void_p(16)  # never dereferenced: validation fails first
    st = L.ccpi_pb_f2D(_dp(h), _dp(v), _dp(ang), 2, len(h), 1, _dp(o), _dp(s), 2, 2, 1,
                       dummy, dummy, None)
    assert st == -2
    assert L.ccpi_last_error()


def test_cone_geometry_rejected():
    """A detector row at the source height (the reference's 'Oops',
    c2D.cpp:487-492), nz < 2 (b2D divides by nz - 1, c2D.cpp:808), a source
    inside the rotation circle, and a source outside the volume's z range,
    where the oracle (the reference's f2D) returns sums off the exact chords
    by orders of magnitude."""
    L = _lib.lib()
    h = np.array([-1.0, 1.0])
    ang = np.array([0.0])
    o = np.array([-2.0, -2.0, -2.0])
    s = np.array([1.0, 1.0, 1.0])
    dummy = C.c_void_p(16)
    v = np.array([-1.0, 0.0, 1.0])
    assert L.ccpi_cb_f2D(-50.0, 0.0, 0.0, 50.0, _dp(h), _dp(v), _dp(ang), dummy, dummy,
                         1, 2, 3, _dp(o), _dp(s), 4, 4, 4, None) == -2
    v = np.array([-1.0, 1.0])
    assert L.ccpi_cb_b2D(-50.0, 0.0, 0.0, 50.0, _dp(h), _dp(v), _dp(ang), dummy, dummy,
                         1, 2, 2, _dp(o), _dp(s), 4, 4, 1, None) == -2
    # source inside the volume's rotation circle
    assert L.ccpi_cb_f2D(-2.0, 0.0, 0.0, 50.0, _dp(h), _dp(v), _dp(ang), dummy, dummy,
                         1, 2, 2, _dp(o), _dp(s), 4, 4, 4, None) == -2
    # source above the volume's z range [-2, 2] (the reference's slice
    # logic has no defined result there)
    assert L.ccpi_cb_b2D(-50.0, 0.0, 2.5, 50.0, _dp(h), _dp(v), _dp(ang), dummy, dummy,
                         1, 2, 2, _dp(o), _dp(s), 4, 4, 4, None) == -2
    assert b"z range" in L.ccpi_last_error()


def test_regularize_rejects_bad_arguments():
    """Mirrored neighbours need every extent >= 2; unknown kinds and
    in-place calls are refused before any device work."""
    L = _lib.lib()
    a, b = C.c_void_p(16), C.c_void_p(32)
    assert L.ccpi_regularize(_lib.CCPI_REG_TV, a, b, 4, 4, 1, None) == -1
    assert L.ccpi_regularize(_lib.CCPI_REG_TIKHONOV, a, b, 1, 4, 4, None) == -1
    assert L.ccpi_regularize(7, a, b, 4, 4, 4, None) == -1
    assert L.ccpi_regularize(_lib.CCPI_REG_TV, a, a, 4, 4, 4, None) == -1
    assert L.ccpi_regularize(_lib.CCPI_REG_TV, None, b, 4, 4, 4, None) == -1
    assert b"extent" in L.ccpi_last_error() or b"null" in L.ccpi_last_error()


def test_cgls_reg_rejects_unknown_kind():
    L = _lib.lib()
    dummy = C.c_void_p(16)
    assert L.ccpi_cgls_reg(None, dummy, dummy, 1, 5, 0.1, None, 0, None) == -1
    assert L.ccpi_cgls_reg(None, dummy, dummy, -1, 0, 0.1, None, 0, None) == -1


def test_find_centre_rejects_bad_arguments():
    """Null geometry, a parallel-beam geometry and a row outside the
    detector are refused before any device work."""
    L = _lib.lib()
    y = C.c_double()
    dummy = C.c_void_p(16)
    assert L.ccpi_xtek_find_centre(None, dummy, 0, C.byref(y), None) == -1
    ang = np.array([0.0, 0.5])
    h = np.array([-1.0, 0.0, 1.0])
    v = np.array([-0.5, 0.5])
    g = _lib.CcpiGeometry(kind=_lib.CCPI_PARALLEL, n_angles=2, n_h=3, n_v=2,
                          angles=_dp(ang), h_pixels=_dp(h), v_pixels=_dp(v),
                          nx=2, ny=2, nz=2)
    g.voxel_size[:] = [1.0, 1.0, 1.0]
    g.origin[:] = [-1.0, -1.0, -1.0]
    assert L.ccpi_xtek_find_centre(C.byref(g), dummy, 0, C.byref(y), None) == -1
    assert b"cone" in L.ccpi_last_error()
    g.kind = _lib.CCPI_CONE
    g.source[:] = [-50.0, 0.0, 0.0]
    g.detector_x = 50.0
    assert L.ccpi_xtek_find_centre(C.byref(g), dummy, 5, C.byref(y), None) == -1
    assert b"v_slice" in L.ccpi_last_error()


def test_noalloc_fp_rejects_tables_beyond_its_scratch():
    """The allocation-free parallel FP holds a geometry's tables in a fixed
    1 MiB device array: 40 bytes per angle alone exceed it at 30000 angles,
    which it rejects before any device work."""
    L = _lib.lib()
    na = 30000
    h = np.linspace(-4, 4, 9)
    v = np.array([0.5])
    ang = np.linspace(0, np.pi, na, endpoint=False)
    o = np.array([-4.0, -4.0, 0.0])
    s = np.array([1.0, 1.0, 1.0])
    dummy = C.c_void_p(16)  # never dereferenced: the size check fails first
    st = L.ccpi_pb_f2D_noalloc(_dp(h), _dp(v), _dp(ang), na, len(h), 1, _dp(o), _dp(s), 8, 8, 1,
                               dummy, dummy, None)
    assert st == -5
    assert b"scratch" in L.ccpi_last_error()
