"""ctypes front end of the CPU oracle (``liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product
path (``ccpi-reconstruction_amd``) never imports it.

The C restatement follows the reference's dispatched projectors and
iteration bodies (see ``oracle_impl.h`` for the file:line map):
``pb_f2D``/``pb_b2D`` restate Core/src/p2D.cpp:468-611 / 950-1087,
``cb_f2D``/``cb_b2D`` restate Core/src/c2D.cpp:469-625 / 948-1119,
``cgls``/``sirt``/``mlem`` restate Core/src/Algorithms/{cgls,sirt,mlem}.cpp.

Every entry point exists in a float flavour (``dtype=np.float32``, the
reference's own precision choices) and a double flavour (``np.float64``,
the FP64 oracle of the same discrete model).
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> str:
    """Compile liboracle.so with the reference's compiler flags."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)
    return _LIB_PATH


class _Geom(C.Structure):
    _fields_ = [
        ("cone", C.c_int),
        ("na", C.c_int), ("nh", C.c_int), ("nv", C.c_int),
        ("angles", C.POINTER(C.c_double)),
        ("h_pix", C.POINTER(C.c_double)),
        ("v_pix", C.POINTER(C.c_double)),
        ("sx", C.c_double), ("sy", C.c_double), ("sz", C.c_double),
        ("det_x", C.c_double),
        ("origin", C.c_double * 3), ("vsize", C.c_double * 3),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        for sfx, ct in (("f32", C.c_float), ("f64", C.c_double)):
            tp = C.POINTER(ct)
            for name in ("or_pb_f2D", "or_pb_b2D"):
                f = getattr(_lib, f"{name}_{sfx}")
                f.argtypes = [dp, C.c_int, dp, C.c_int, dp, C.c_int, dp, dp,
                              C.c_int, C.c_int, C.c_int, tp, tp]
                f.restype = None
            for name in ("or_cb_f2D", "or_cb_b2D"):
                f = getattr(_lib, f"{name}_{sfx}")
                f.argtypes = [C.c_double] * 4 + [dp, C.c_int, dp, C.c_int, dp,
                                                 C.c_int, dp, dp, C.c_int,
                                                 C.c_int, C.c_int, tp, tp]
                f.restype = None
            f = getattr(_lib, f"or_cgls_{sfx}")
            f.argtypes = [C.POINTER(_Geom), tp, tp, C.c_int, dp]
            f.restype = C.c_int
            f = getattr(_lib, f"or_cgls_track_{sfx}")
            f.argtypes = [C.POINTER(_Geom), tp, tp, C.c_int, dp, C.POINTER(C.c_float), dp, dp]
            f.restype = C.c_int
            for name in ("or_sirt", "or_mlem"):
                f = getattr(_lib, f"{name}_{sfx}")
                f.argtypes = [C.POINTER(_Geom), tp, tp, C.c_int]
                f.restype = C.c_int
            for name in ("or_tikhonov", "or_tv"):
                f = getattr(_lib, f"{name}_{sfx}")
                f.argtypes = [tp, tp, C.c_int, C.c_int, C.c_int]
                f.restype = None
            f = getattr(_lib, f"or_cgls_reg_{sfx}")
            _lib.or_find_centre.argtypes = [C.POINTER(_Geom), C.POINTER(C.c_float), C.c_int,
                                            C.POINTER(C.c_double), C.POINTER(C.c_double)]
            _lib.or_find_centre.restype = C.c_int
            f.argtypes = [C.POINTER(_Geom), tp, tp, C.c_int, C.c_int, C.c_double, dp]
            f.restype = C.c_int
    return _lib


def _sfx(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return "f32", C.c_float
    if dtype == np.float64:
        return "f64", C.c_double
    raise TypeError(f"oracle supports float32/float64, not {dtype}")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _check(a, shape, dtype, name):
    if a.shape != tuple(shape) or a.dtype != np.dtype(dtype) or not a.flags.c_contiguous:
        raise ValueError(f"{name}: need C-contiguous {np.dtype(dtype)} {tuple(shape)},"
                         f" got {a.dtype} {a.shape}")


def pb_f2D(h_pix, v_pix, angles, origin, vsize, vol, out=None):
    """Parallel-beam forward projection; ``out`` [na][nh][nv] is accumulated."""
    h, v, a = _d(h_pix), _d(v_pix), _d(angles)
    o, s = _d(origin), _d(vsize)
    nx, ny, nz = vol.shape
    sfx, ct = _sfx(vol.dtype)
    if out is None:
        out = np.zeros((a.size, h.size, v.size), vol.dtype)
    _check(vol, (nx, ny, nz), vol.dtype, "vol")
    _check(out, (a.size, h.size, v.size), vol.dtype, "out")
    getattr(lib(), f"or_pb_f2D_{sfx}")(_dp(h), h.size, _dp(v), v.size, _dp(a), a.size,
                                       _dp(o), _dp(s), nx, ny, nz, _ptr(vol, ct), _ptr(out, ct))
    return out


def pb_b2D(h_pix, v_pix, angles, origin, vsize, sino, shape, out=None):
    """Parallel-beam backprojection; ``out`` [nx][ny][nz] is accumulated."""
    h, v, a = _d(h_pix), _d(v_pix), _d(angles)
    o, s = _d(origin), _d(vsize)
    nx, ny, nz = shape
    sfx, ct = _sfx(sino.dtype)
    if out is None:
        out = np.zeros((nx, ny, nz), sino.dtype)
    _check(sino, (a.size, h.size, v.size), sino.dtype, "sino")
    _check(out, (nx, ny, nz), sino.dtype, "out")
    getattr(lib(), f"or_pb_b2D_{sfx}")(_dp(h), h.size, _dp(v), v.size, _dp(a), a.size,
                                       _dp(o), _dp(s), nx, ny, nz, _ptr(sino, ct), _ptr(out, ct))
    return out


def cb_f2D(source, det_x, h_pix, v_pix, angles, origin, vsize, vol, out=None):
    """Cone-beam forward projection: ``out = (out + A vol) * d_conv``."""
    h, v, a = _d(h_pix), _d(v_pix), _d(angles)
    o, s = _d(origin), _d(vsize)
    nx, ny, nz = vol.shape
    sfx, ct = _sfx(vol.dtype)
    if out is None:
        out = np.zeros((a.size, h.size, v.size), vol.dtype)
    _check(vol, (nx, ny, nz), vol.dtype, "vol")
    _check(out, (a.size, h.size, v.size), vol.dtype, "out")
    sx, sy, sz = (float(q) for q in source)
    getattr(lib(), f"or_cb_f2D_{sfx}")(sx, sy, sz, float(det_x), _dp(h), h.size, _dp(v), v.size,
                                       _dp(a), a.size, _dp(o), _dp(s), nx, ny, nz,
                                       _ptr(vol, ct), _ptr(out, ct))
    return out


def cb_b2D(source, det_x, h_pix, v_pix, angles, origin, vsize, sino, shape, out=None):
    """Cone-beam backprojection: ``out += A^T (sino * d_conv)``."""
    h, v, a = _d(h_pix), _d(v_pix), _d(angles)
    o, s = _d(origin), _d(vsize)
    nx, ny, nz = shape
    sfx, ct = _sfx(sino.dtype)
    if out is None:
        out = np.zeros((nx, ny, nz), sino.dtype)
    _check(sino, (a.size, h.size, v.size), sino.dtype, "sino")
    _check(out, (nx, ny, nz), sino.dtype, "out")
    sx, sy, sz = (float(q) for q in source)
    getattr(lib(), f"or_cb_b2D_{sfx}")(sx, sy, sz, float(det_x), _dp(h), h.size, _dp(v), v.size,
                                       _dp(a), a.size, _dp(o), _dp(s), nx, ny, nz,
                                       _ptr(sino, ct), _ptr(out, ct))
    return out


class Geometry:
    """Plain description of one projector instance (host arrays)."""

    def __init__(self, angles, h_pix, v_pix, origin, vsize, shape,
                 cone=False, source=(0.0, 0.0, 0.0), det_x=0.0):
        self.angles = _d(angles)
        self.h_pix = _d(h_pix)
        self.v_pix = _d(v_pix)
        self.origin = _d(origin)
        self.vsize = _d(vsize)
        self.shape = tuple(int(q) for q in shape)
        self.cone = bool(cone)
        self.source = tuple(float(q) for q in source)
        self.det_x = float(det_x)

    @property
    def sino_shape(self):
        return (self.angles.size, self.h_pix.size, self.v_pix.size)

    def fp(self, vol, out=None):
        if self.cone:
            return cb_f2D(self.source, self.det_x, self.h_pix, self.v_pix, self.angles,
                          self.origin, self.vsize, vol, out)
        return pb_f2D(self.h_pix, self.v_pix, self.angles, self.origin, self.vsize, vol, out)

    def bp(self, sino, out=None):
        if self.cone:
            return cb_b2D(self.source, self.det_x, self.h_pix, self.v_pix, self.angles,
                          self.origin, self.vsize, sino, self.shape, out)
        return pb_b2D(self.h_pix, self.v_pix, self.angles, self.origin, self.vsize,
                      sino, self.shape, out)

    def _struct(self):
        g = _Geom()
        g.cone = int(self.cone)
        g.na, g.nh, g.nv = self.sino_shape
        g.angles, g.h_pix, g.v_pix = _dp(self.angles), _dp(self.h_pix), _dp(self.v_pix)
        g.sx, g.sy, g.sz = self.source
        g.det_x = self.det_x
        g.origin[:] = list(self.origin)
        g.vsize[:] = list(self.vsize)
        g.nx, g.ny, g.nz = self.shape
        return g


def cgls(geom: Geometry, sino, niter, x=None, dtype=np.float32):
    """CGLS body; returns (x, residual, norms) with norms = ||A^T r|| per iter."""
    sfx, ct = _sfx(dtype)
    b = np.array(sino, dtype=dtype, order="C", copy=True)
    x = np.zeros(geom.shape, dtype) if x is None else np.array(x, dtype=dtype, order="C")
    norms = np.zeros(max(niter, 1), np.float64)
    g = geom._struct()
    getattr(lib(), f"or_cgls_{sfx}")(C.byref(g), _ptr(b, ct), _ptr(x, ct), niter, _dp(norms))
    return x, b, norms[:niter]


def cgls_track(geom: Geometry, sino, niter, truth=None, dtype=np.float32, timing=False):
    """CGLS body from x = 0 that also records RMSE(x - truth) after every
    iteration (truth given) and each iteration's wall time; returns (x,
    residual, norms, rmse, seconds per iteration)."""
    sfx, ct = _sfx(dtype)
    b = np.array(sino, dtype=dtype, order="C", copy=True)
    x = np.zeros(geom.shape, dtype)
    norms = np.zeros(max(niter, 1), np.float64)
    rmse = np.zeros(max(niter, 1), np.float64)
    secs = np.zeros(max(niter, 1), np.float64)
    tp = None
    if truth is not None:
        t = np.ascontiguousarray(truth, dtype=np.float32)
        tp = _ptr(t, C.c_float)
    g = geom._struct()
    getattr(lib(), f"or_cgls_track_{sfx}")(C.byref(g), _ptr(b, ct), _ptr(x, ct), niter,
                                           _dp(norms), tp, _dp(rmse), _dp(secs))
    return x, b, norms[:niter], rmse[:niter], secs[:niter]


def sirt(geom: Geometry, sino, niter, x=None, dtype=np.float32):
    sfx, ct = _sfx(dtype)
    b = np.ascontiguousarray(sino, dtype=dtype)
    x = np.zeros(geom.shape, dtype) if x is None else np.array(x, dtype=dtype, order="C")
    g = geom._struct()
    getattr(lib(), f"or_sirt_{sfx}")(C.byref(g), _ptr(b, ct), _ptr(x, ct), niter)
    return x


def mlem(geom: Geometry, sino, niter, dtype=np.float32):
    sfx, ct = _sfx(dtype)
    b = np.ascontiguousarray(sino, dtype=dtype)
    x = np.ones(geom.shape, dtype)
    g = geom._struct()
    getattr(lib(), f"or_mlem_{sfx}")(C.byref(g), _ptr(b, ct), _ptr(x, ct), niter)
    return x


TIKHONOV, TV = 0, 1


def regularize(kind: int, vol, dtype=np.float32):
    """Tikhonov (kind 0, Core/src/tikhonov.cpp:4-32) or TV (kind 1,
    Core/src/total_v.cpp:21-195) term of a [nx][ny][nz] volume."""
    sfx, ct = _sfx(dtype)
    a = np.ascontiguousarray(vol, dtype=dtype)
    out = np.empty_like(a)
    nx, ny, nz = a.shape
    name = "or_tikhonov" if kind == TIKHONOV else "or_tv"
    getattr(lib(), f"{name}_{sfx}")(_ptr(a, ct), _ptr(out, ct), nx, ny, nz)
    return out


def cgls_reg(geom: Geometry, sino, niter, kind: int, lam: float, x=None, dtype=np.float32):
    """Regularised CGLS body (Core/src/Algorithms/cgls.cpp:343-435); returns
    (x, residual, norms)."""
    sfx, ct = _sfx(dtype)
    b = np.array(sino, dtype=dtype, order="C", copy=True)
    x = np.zeros(geom.shape, dtype) if x is None else np.array(x, dtype=dtype, order="C")
    norms = np.zeros(max(niter, 1), np.float64)
    g = geom._struct()
    getattr(lib(), f"or_cgls_reg_{sfx}")(C.byref(g), _ptr(b, ct), _ptr(x, ct), niter, kind,
                                         float(lam), _dp(norms))
    return x, b, norms[:niter]


def find_centre(geom: Geometry, pixels, v_slice: int):
    """XTek centre-of-rotation search (Core/src/Readers/xtek.cpp:671-758) on
    row v_slice of a [angle][h][v] float sinogram.  Returns (y_centre,
    M[5][21]), or (None, M) when the search fails."""
    p = np.ascontiguousarray(pixels, dtype=np.float32)
    if p.shape != geom.sino_shape:
        raise ValueError(f"pixels must be {geom.sino_shape}")
    g = geom._struct()
    y = C.c_double()
    m = np.zeros((5, 21), np.float64)
    st = lib().or_find_centre(C.byref(g), p.ctypes.data_as(C.POINTER(C.c_float)), int(v_slice),
                              C.byref(y), _dp(m))
    return (None if st else y.value), m
