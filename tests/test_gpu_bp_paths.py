"""Every cone backprojection kernel instance against the oracle (GPU test).

The cone BP picks its row stride S from the largest number of detector rows
per slice (at least 1.05 slices between the rows of one slot) and falls back
to LDS atomics on one array when S would exceed 4 or the S arrays would not
fit 64 KB per workgroup.  These geometries select S = 1, 2 (the default test
cases), 3, 4 and both atomic fallbacks; each is checked for FP and BP parity
(1e-6 / 1e-5 of the maximum, as in test_gpu_parity) and the adjoint
identity.  ``_path`` restates the host's choice so that a geometry that no
longer reaches its intended kernel fails here instead of silently testing
another one.
"""

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle
from geometries import CB, centred, ellipsoid_phantom, oracle_geom

pytestmark = pytest.mark.gpu

CASES = {
    "s1": (CB((10, 10, 8), (1, 1, 1), np.linspace(0, 2 * np.pi, 12, endpoint=False) + 0.1,
              centred(20, 1.6), centred(8, 3.0), 40.0, 80.0), (1, False)),
    "s3": (CB((10, 10, 6), (1, 1, 1), np.linspace(0, 2 * np.pi, 12, endpoint=False) + 0.1,
              centred(20, 1.6), centred(20, 0.95), 40.0, 80.0), (3, False)),
    "s4": (CB((10, 10, 6), (1, 1, 1), np.linspace(0, 2 * np.pi, 12, endpoint=False) + 0.1,
              centred(20, 1.6), centred(24, 0.75), 40.0, 80.0), (4, False)),
    "atomic_stride": (CB((8, 8, 4), (1, 1, 1), np.linspace(0, 2 * np.pi, 10, endpoint=False),
                         centred(16, 1.6), centred(40, 0.4), 40.0, 80.0), (7, True)),
    "atomic_lds": (CB((4, 4, 1500), (1, 1, 0.05), np.linspace(0, 2 * np.pi, 8, endpoint=False),
                      centred(10, 1.2), centred(3000, 0.05), 40.0, 80.0), (3, True)),
}


def _path(g):
    """(row stride, atomic) as the host chooses them (cone.hip plan build)."""
    nx, ny, nz = g.shape
    vs, o = g.vsize, g.origin
    R = max(math.hypot(o[0] + cx * nx * vs[0], o[1] + cy * ny * vs[1])
            for cx in (0, 1) for cy in (0, 1))
    Lmax = max(math.hypot(g.det_x - g.source[0], g.h_pix[i] - g.source[1]) for i in (0, -1))
    amin = max((math.hypot(g.source[0], g.source[1]) - R) / Lmax, 1e-3)
    want = max(1, math.ceil(vs[2] / (amin * np.diff(g.v_pix).min()) * 1.05))
    s = min(16, want)
    pad = (nz + 2 + 3) & ~3
    lds = 4 * (pad + 4 * s * pad)
    return s, not (want <= 16 and s <= 4 and lds <= 64 * 1024)


def _geom(g):
    from ccpi.reconstruction.projectors import Geometry
    return Geometry(g.angles, g.h_pix, g.v_pix, g.origin, g.vsize, g.shape, cone=True,
                    source=g.source, detector_x=g.det_x)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("name", list(CASES))
def test_bp_path_parity(name):
    from ccpi.reconstruction.projectors import backward_project, forward_project
    g, want_path = CASES[name]
    assert _path(g) == want_path
    og = oracle_geom(g)
    x = ellipsoid_phantom(g.shape, g.vsize, g.origin, 4, seed=7).astype(np.float32)
    print(f"\ncone BP path {name}: S={want_path[0]} atomic={want_path[1]}", flush=True)
    y = forward_project(_geom(g), torch.from_numpy(x).cuda()).cpu().numpy()
    assert _rel(y, og.fp(x)) < 1e-6
    s = np.random.default_rng(8).random(g.sino_shape).astype(np.float32)
    b = backward_project(_geom(g), torch.from_numpy(s).cuda()).cpu().numpy()
    assert _rel(b, og.bp(s)) < 1e-5
    lhs = float(np.vdot(y.astype(np.float64), s))
    rhs = float(np.vdot(x.astype(np.float64), b))
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)
