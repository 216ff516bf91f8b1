"""Projectors and iteration bodies on device-resident torch tensors.

Thin host layer over the C ABI (include/ccpi_hip.h): tensors stay in HBM,
calls are enqueued on torch's current stream.  torch provides the device
memory and the stream; all arithmetic happens in the HIP kernels of
libccpi_hip.so.

Layouts (Core/src/base_types.hpp:60-63): volume [nx][ny][nz] (z fastest),
sinogram [n_angles][n_h][n_v] (v fastest), both float32.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from ._lib import (CCPI_ACCUM_FP64, CCPI_CONE, CCPI_PARALLEL, CCPI_REDUCE_FP32,
                   CCPI_REG_TIKHONOV, CCPI_REG_TV, CCPI_TIMING, CcpiGeometry, check, lib)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


@dataclass
class Geometry:
    """One projector instance: the state of CCPi::parallel_beam /
    CCPi::cone_beam (Core/src/instruments.hpp:130-143, 194-199).

    angles in radians; h_pixels / v_pixels / origin / voxel_size in mm;
    shape = (nx, ny, nz).  For the cone beam, ``source`` is the source
    position and ``detector_x`` the detector plane at angle 0."""

    angles: np.ndarray
    h_pixels: np.ndarray
    v_pixels: np.ndarray
    origin: Sequence[float]
    voxel_size: Sequence[float]
    shape: Sequence[int]
    cone: bool = False
    source: Sequence[float] = (0.0, 0.0, 0.0)
    detector_x: float = 0.0
    _c: CcpiGeometry | None = field(default=None, repr=False, compare=False)

    def __post_init__(self):
        self.angles = _f64(self.angles).ravel()
        self.h_pixels = _f64(self.h_pixels).ravel()
        self.v_pixels = _f64(self.v_pixels).ravel()
        self.origin = tuple(float(q) for q in self.origin)
        self.voxel_size = tuple(float(q) for q in self.voxel_size)
        self.shape = tuple(int(q) for q in self.shape)
        self.source = tuple(float(q) for q in self.source)
        self.detector_x = float(self.detector_x)

    @property
    def sino_shape(self):
        return (self.angles.size, self.h_pixels.size, self.v_pixels.size)

    def c_struct(self) -> CcpiGeometry:
        if self._c is None:
            g = CcpiGeometry()
            g.kind = CCPI_CONE if self.cone else CCPI_PARALLEL
            g.n_angles, g.n_h, g.n_v = self.sino_shape
            dp = C.POINTER(C.c_double)
            g.angles = self.angles.ctypes.data_as(dp)
            g.h_pixels = self.h_pixels.ctypes.data_as(dp)
            g.v_pixels = self.v_pixels.ctypes.data_as(dp)
            g.source[:] = list(self.source)
            g.detector_x = self.detector_x
            g.origin[:] = list(self.origin)
            g.voxel_size[:] = list(self.voxel_size)
            g.nx, g.ny, g.nz = self.shape
            self._c = g
        return self._c


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(t, shape, name):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name}: need a HIP device tensor")
    if t.dtype != torch.float32 or not t.is_contiguous() or tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: need contiguous float32 {tuple(shape)}, got "
                         f"{t.dtype} {tuple(t.shape)}")
    return C.c_void_p(t.data_ptr())


def forward_project(geom: Geometry, vol, out=None):
    """Sinogram of ``vol``.  With ``out`` given the projection accumulates
    into it (parallel) or into it and then scales by d_conv (cone), as
    f2D does (p2D.cpp:468-611, c2D.cpp:469-625)."""
    import torch
    if out is None:
        out = torch.zeros(geom.sino_shape, dtype=torch.float32, device=vol.device)
    pv = _dev(vol, geom.shape, "vol")
    po = _dev(out, geom.sino_shape, "out")
    check(lib().ccpi_forward_project(C.byref(geom.c_struct()), po, pv, _stream()))
    return out


def backward_project(geom: Geometry, sino, out=None):
    """Backprojection of ``sino``, accumulated into ``out`` if given
    (p2D.cpp:950-1087, c2D.cpp:948-1119)."""
    import torch
    if out is None:
        out = torch.zeros(geom.shape, dtype=torch.float32, device=sino.device)
    ps = _dev(sino, geom.sino_shape, "sino")
    po = _dev(out, geom.shape, "out")
    check(lib().ccpi_backward_project(C.byref(geom.c_struct()), ps, po, _stream()))
    return out


def cgls(geom: Geometry, b, iterations: int, x=None, reduce_fp32: bool = False):
    """CGLS body (cgls.cpp:26-157).  Returns (x, residual, norms) where the
    residual is b - A x and norms[i] = ||A^T r|| after iteration i.  ``b``
    is not modified (a copy carries the residual)."""
    import torch
    r = b.detach().clone().contiguous()
    if x is None:
        x = torch.zeros(geom.shape, dtype=torch.float32, device=b.device)
    norms = np.zeros(max(iterations, 1), np.float64)
    check(lib().ccpi_cgls(C.byref(geom.c_struct()), _dev(r, geom.sino_shape, "b"),
                          _dev(x, geom.shape, "x"), int(iterations),
                          norms.ctypes.data_as(C.POINTER(C.c_double)),
                          CCPI_REDUCE_FP32 if reduce_fp32 else 0, _stream()))
    return x, r, norms[:iterations]


TIKHONOV, TV = CCPI_REG_TIKHONOV, CCPI_REG_TV


def regularize(kind: int, vol, out=None):
    """Tikhonov (``TIKHONOV``, tikhonov.cpp:4-32) or TV (``TV``,
    total_v.cpp:21-195) term of a [nx][ny][nz] float32 device volume; every
    extent must be >= 2."""
    import torch
    if vol.dim() != 3:
        raise ValueError("vol must be [nx][ny][nz]")
    shape = tuple(vol.shape)
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=vol.device)
    check(lib().ccpi_regularize(int(kind), _dev(vol, shape, "vol"), _dev(out, shape, "out"),
                                shape[0], shape[1], shape[2], _stream()))
    return out


def cgls_reg(geom: Geometry, b, iterations: int, kind: int, lam: float, x=None,
             reduce_fp32: bool = False):
    """Regularised CGLS, cgls_tikhonov / cgls_tv_reg (cgls.cpp:343-435).
    Returns (x, residual, norms) like :func:`cgls`; ``b`` is not modified."""
    import torch
    r = b.detach().clone().contiguous()
    if x is None:
        x = torch.zeros(geom.shape, dtype=torch.float32, device=b.device)
    norms = np.zeros(max(iterations, 1), np.float64)
    check(lib().ccpi_cgls_reg(C.byref(geom.c_struct()), _dev(r, geom.sino_shape, "b"),
                              _dev(x, geom.shape, "x"), int(iterations), int(kind),
                              float(lam), norms.ctypes.data_as(C.POINTER(C.c_double)),
                              CCPI_REDUCE_FP32 if reduce_fp32 else 0, _stream()))
    return x, r, norms[:iterations]


def sirt(geom: Geometry, b, iterations: int, x=None):
    """SIRT body (sirt.cpp:15-80), starting from ``x`` (zeros if None)."""
    import torch
    if x is None:
        x = torch.zeros(geom.shape, dtype=torch.float32, device=b.device)
    check(lib().ccpi_sirt(C.byref(geom.c_struct()), _dev(b, geom.sino_shape, "b"),
                          _dev(x, geom.shape, "x"), int(iterations), 0, _stream()))
    return x


def mlem(geom: Geometry, b, iterations: int, x=None, accum_fp64: bool = False):
    """MLEM body (mlem.cpp:15-80); x is reset to ones first.  With
    ``accum_fp64`` the projections' sums (A x, A^T (b / A x), A^T 1) are
    accumulated in double (C ABI flag CCPI_ACCUM_FP64)."""
    import torch
    if x is None:
        x = torch.empty(geom.shape, dtype=torch.float32, device=b.device)
    check(lib().ccpi_mlem(C.byref(geom.c_struct()), _dev(b, geom.sino_shape, "b"),
                          _dev(x, geom.shape, "x"), int(iterations),
                          CCPI_ACCUM_FP64 if accum_fp64 else 0, _stream()))
    return x


def cgls_batched(geoms: Sequence[Geometry], bs, iterations: int, xs=None,
                 reduce_fp32: bool = False):
    """Independent parallel-beam CGLS problems of uneven sizes in one call.
    Returns (xs, residuals, norms[count][iterations])."""
    import torch
    n = len(geoms)
    rs = [b.detach().clone().contiguous() for b in bs]
    if xs is None:
        xs = [torch.zeros(g.shape, dtype=torch.float32, device=b.device)
              for g, b in zip(geoms, bs)]
    arr = (CcpiGeometry * n)(*[g.c_struct() for g in geoms])
    pb = (C.c_void_p * n)(*[_dev(r, g.sino_shape, "b").value for r, g in zip(rs, geoms)])
    px = (C.c_void_p * n)(*[_dev(x, g.shape, "x").value for x, g in zip(xs, geoms)])
    norms = np.zeros((n, max(iterations, 1)), np.float64)
    check(lib().ccpi_pb_cgls_batched(arr, n, pb, px, int(iterations),
                                     norms.ctypes.data_as(C.POINTER(C.c_double)),
                                     CCPI_REDUCE_FP32 if reduce_fp32 else 0, _stream()))
    return xs, rs, norms[:, :iterations]


class CglsSession:
    """CGLS split into launch-only steps (C ABI ccpi_cgls_create/iterate):
    iterate() enqueues work on the current stream without host sync.  The
    session holds ``b`` (overwritten with the residual) and ``x``."""

    def __init__(self, geom: Geometry, b, x, max_iterations: int, reduce_fp32=False,
                 timing=False):
        self.geom, self.b, self.x = geom, b, x
        flags = (CCPI_REDUCE_FP32 if reduce_fp32 else 0) | (CCPI_TIMING if timing else 0)
        h = C.c_void_p()
        check(lib().ccpi_cgls_create(C.byref(geom.c_struct()), _dev(b, geom.sino_shape, "b"),
                                     _dev(x, geom.shape, "x"), int(max_iterations), flags,
                                     _stream(), C.byref(h)))
        self._h = h

    def iterate(self, n: int = 1) -> None:
        check(lib().ccpi_cgls_iterate(self._h, int(n)))

    def norms(self) -> np.ndarray:
        buf = np.zeros(4096, np.float64)
        done = C.c_int()
        check(lib().ccpi_cgls_norms(self._h, buf.ctypes.data_as(C.POINTER(C.c_double)),
                                    C.byref(done)))
        return buf[:done.value].copy()

    def timings(self):
        """(fp_ms_total, fp_count, bp_ms_total, bp_count) from HIP events."""
        fpm, bpm = C.c_double(), C.c_double()
        fpc, bpc = C.c_int(), C.c_int()
        check(lib().ccpi_cgls_timings(self._h, C.byref(fpm), C.byref(fpc), C.byref(bpm),
                                      C.byref(bpc)))
        return fpm.value, fpc.value, bpm.value, bpc.value

    def close(self) -> None:
        if self._h:
            check(lib().ccpi_cgls_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def projector_nnz(geom: Geometry) -> int:
    """Nonzero system-matrix entries of the projector, counted on device."""
    n = C.c_ulonglong()
    check(lib().ccpi_projector_nnz(C.byref(geom.c_struct()), C.byref(n), _stream()))
    return int(n.value)


def plan_cache_clear() -> None:
    lib().ccpi_plan_cache_clear()


def plan_cache_bytes() -> int:
    return int(lib().ccpi_plan_cache_bytes())
