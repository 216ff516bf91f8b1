"""Every cone backprojection kernel instance against the oracle (GPU test).

The cone BP picks its row stride S from the largest number of detector rows
per slice (at least 1.05 slices between the rows of one slot) and uses S
LDS arrays per wavefront while S <= 4 and they fit 64 KB per workgroup.
Otherwise it takes the phased form: S = 4, 8, 12 or 16 slots on one array,
written slot after slot, and within a slot the lanes in ceil(want / S)
phases (past 4 rows per slice, S = 8, 12 or 16 by the least padded rows x
(1 + 0.11 per extra phase)); each column runs its own phases, from the
smallest walk parameter of its rays, at most the layout's.  These
geometries select S = 1, 2 (the default test cases), 3, 4, and S = 2 over
three 128-row blocks (one under the source row, one across it, one above it
with padded rows, so the row-set addresses of both directions and of the
mixed block run, with floor-form and main rows in one block), and the phased
form with:
  * S = 8: fewer than 64 rows, with the source at mid-height and near the
    volume's top; a padded block of main rows (every ray inside the
    volume's z range); two blocks, the second padded, with rows outside it;
  * S = 8 in two phases (300 rows); S = 12 in one and in two phases (700
    rows: one block of 768), S = 16 in one and in two phases (1000 rows);
    the wide cones' corners close to the source, so columns near the axis
    run fewer phases than the corners;
  * S = 4 (arrays beyond the LDS budget);
  * S = 12 in 64 phases: the source 0.002 mm outside the volume's corner
    circle, where thousands of rows fall on one slice near the corners.
Each is checked for FP and BP parity (1e-6 / 1e-5 of the maximum, as in
test_gpu_parity) and the adjoint identity.  ``_path`` restates the host's
choice so that a geometry that no longer reaches its intended kernel fails
here instead of silently testing another one.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from bp_cases import CASES, NOT_ADJOINT, bp_path as _path
from geometries import ellipsoid_phantom, oracle_geom

pytestmark = pytest.mark.gpu


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
    print(f"\ncone BP path {name}: S={want_path[0]} phased={want_path[1]} "
          f"phases={want_path[2]}", flush=True)
    y = forward_project(_geom(g), torch.from_numpy(x).cuda()).cpu().numpy()
    assert _rel(y, og.fp(x)) < 1e-6
    s = np.random.default_rng(8).random(g.sino_shape).astype(np.float32)
    b = backward_project(_geom(g), torch.from_numpy(s).cuda()).cpu().numpy()
    assert _rel(b, og.bp(s)) < 1e-5
    if name in NOT_ADJOINT:
        return
    lhs = float(np.vdot(y.astype(np.float64), s))
    rhs = float(np.vdot(x.astype(np.float64), b))
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)
