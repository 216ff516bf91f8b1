"""ctypes binding of libccpi_hip.so (C ABI in include/ccpi_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
ccpi-reconstruction_amd/csrc``).  Nothing here falls back to a CPU path:
loading fails loudly when the library is missing.
"""

from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libccpi_hip.so")

CCPI_PARALLEL = 0
CCPI_CONE = 1
CCPI_REDUCE_FP32 = 1
CCPI_TIMING = 2
CCPI_ACCUM_FP64 = 4
CCPI_REG_TIKHONOV = 0
CCPI_REG_TV = 1

# every symbol include/ccpi_hip.h declares
EXPORTS = (
    "ccpi_pb_f2D", "ccpi_pb_f2D_noalloc", "ccpi_pb_b2D", "ccpi_cb_f2D", "ccpi_cb_b2D",
    "ccpi_forward_project", "ccpi_backward_project",
    "ccpi_cgls", "ccpi_sirt", "ccpi_mlem", "ccpi_regularize", "ccpi_cgls_reg",
    "ccpi_xtek_find_centre",
    "ccpi_pb_cgls_batched", "ccpi_cgls_create", "ccpi_cgls_iterate", "ccpi_cgls_norms", "ccpi_cgls_timings",
    "ccpi_cgls_destroy", "ccpi_projector_nnz", "ccpi_cone_bp_layout",
    "ccpi_error_string", "ccpi_last_error", "ccpi_version",
    "ccpi_plan_cache_clear", "ccpi_plan_cache_bytes",
)


class CcpiError(RuntimeError):
    def __init__(self, status: int, detail: str):
        self.status = status
        super().__init__(f"ccpi status {status}: {detail}")


class CcpiGeometry(C.Structure):
    """Mirror of ``ccpi_geometry``."""
    _fields_ = [
        ("kind", C.c_int),
        ("n_angles", C.c_int), ("n_h", C.c_int), ("n_v", C.c_int),
        ("angles", C.POINTER(C.c_double)),
        ("h_pixels", C.POINTER(C.c_double)),
        ("v_pixels", C.POINTER(C.c_double)),
        ("source", C.c_double * 3),
        ("detector_x", C.c_double),
        ("origin", C.c_double * 3),
        ("voxel_size", C.c_double * 3),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
    ]


_lib = None


def lib():
    """Load the HIP library once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP extension first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    dp, fp, vp, i = C.POINTER(C.c_double), C.c_void_p, C.c_void_p, C.c_int
    d = C.c_double
    L.ccpi_pb_f2D.argtypes = [dp, dp, dp, i, i, i, dp, dp, i, i, i, fp, fp, vp]
    L.ccpi_pb_f2D_noalloc.argtypes = [dp, dp, dp, i, i, i, dp, dp, i, i, i, fp, fp, vp]
    L.ccpi_pb_b2D.argtypes = [dp, dp, dp, fp, fp, i, i, i, dp, dp, i, i, i, vp]
    L.ccpi_cb_f2D.argtypes = [d, d, d, d, dp, dp, dp, fp, fp, i, i, i, dp, dp, i, i, i, vp]
    L.ccpi_cb_b2D.argtypes = [d, d, d, d, dp, dp, dp, fp, fp, i, i, i, dp, dp, i, i, i, vp]
    gp = C.POINTER(CcpiGeometry)
    L.ccpi_forward_project.argtypes = [gp, fp, fp, vp]
    L.ccpi_backward_project.argtypes = [gp, fp, fp, vp]
    L.ccpi_cgls.argtypes = [gp, fp, fp, i, dp, i, vp]
    L.ccpi_sirt.argtypes = [gp, fp, fp, i, i, vp]
    L.ccpi_mlem.argtypes = [gp, fp, fp, i, i, vp]
    L.ccpi_regularize.argtypes = [i, fp, fp, i, i, i, vp]
    L.ccpi_cgls_reg.argtypes = [gp, fp, fp, i, i, d, dp, i, vp]
    L.ccpi_xtek_find_centre.argtypes = [gp, fp, i, dp, vp]
    L.ccpi_pb_cgls_batched.argtypes = [gp, i, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                       i, dp, i, vp]
    ip = C.POINTER(C.c_int)
    L.ccpi_cgls_create.argtypes = [gp, fp, fp, i, i, vp, C.POINTER(C.c_void_p)]
    L.ccpi_cgls_iterate.argtypes = [vp, i]
    L.ccpi_cgls_norms.argtypes = [vp, dp, ip]
    L.ccpi_cgls_timings.argtypes = [vp, dp, ip, dp, ip]
    L.ccpi_cgls_destroy.argtypes = [vp]
    L.ccpi_projector_nnz.argtypes = [gp, C.POINTER(C.c_ulonglong), vp]
    L.ccpi_cone_bp_layout.argtypes = [gp, ip, ip, ip, ip]
    for name in ("ccpi_pb_f2D", "ccpi_pb_f2D_noalloc", "ccpi_pb_b2D", "ccpi_cb_f2D", "ccpi_cb_b2D",
                 "ccpi_forward_project", "ccpi_backward_project", "ccpi_cgls",
                 "ccpi_sirt", "ccpi_mlem", "ccpi_regularize", "ccpi_cgls_reg",
                 "ccpi_xtek_find_centre", "ccpi_pb_cgls_batched", "ccpi_cgls_create",
                 "ccpi_cgls_iterate", "ccpi_cgls_norms", "ccpi_cgls_timings",
                 "ccpi_cgls_destroy", "ccpi_projector_nnz", "ccpi_cone_bp_layout"):
        getattr(L, name).restype = C.c_int
    L.ccpi_error_string.argtypes = [i]
    L.ccpi_error_string.restype = C.c_char_p
    L.ccpi_last_error.argtypes = []
    L.ccpi_last_error.restype = C.c_char_p
    L.ccpi_version.restype = C.c_int
    L.ccpi_plan_cache_clear.restype = None
    L.ccpi_plan_cache_bytes.restype = C.c_longlong
    _lib = L
    return L


def check(status: int) -> None:
    if status != 0:
        L = lib()
        detail = (L.ccpi_last_error() or b"").decode() or L.ccpi_error_string(status).decode()
        raise CcpiError(status, detail)
