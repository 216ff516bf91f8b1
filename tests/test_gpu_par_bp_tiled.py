"""The tiled parallel-beam backprojector against the oracle (GPU test).

With one detector row per slice, the parallel BP stages each 16 x 16
column tile's detector columns in LDS per block of angles (16 bytes at a
time when the row length is a multiple of 4, else one float at a time),
8 or 16 slices per lane, and splits the angles into groups (summed through
partials) while the tiles alone would give fewer than 4096 waves.  These
geometries reach: ragged tiles and a partial last slice chunk (nz = 20,
rows beyond the volume), pixels finer than the voxels (wide staged spans),
8 slices per lane with many angle groups, rows of odd length (nz = nv =
7, 13), and one angle group (a large volume).  Each is checked against the
oracle's b2D (1e-5 of the
maximum, the summation-order tolerance of the other BP tests), for the
adjoint identity with the FP, for accumulation into a non-zero volume, and
for run-to-run identity.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle
from geometries import PB, centred

pytestmark = pytest.mark.gpu

CASES = {
    "ragged_chunks": PB((37, 21, 20), (1.0, 1.0, 1.0),
                        np.deg2rad(np.r_[0.0, 45.0, 90.0, np.linspace(3, 177, 27)]),
                        centred(56, 1.0, 0.31), centred(24, 1.0)),
    "fine_pixels": PB((24, 30, 12), (1.0, 1.0, 1.0), np.deg2rad(np.linspace(0, 180, 40,
                                                                           endpoint=False)),
                      centred(110, 0.35, -0.2), centred(12, 1.0)),
    "r8_groups": PB((16, 16, 4), (1.0, 1.0, 1.0), np.deg2rad(np.linspace(0, 180, 64,
                                                                        endpoint=False) + 1),
                    centred(24, 1.0, 0.37), centred(4, 1.0)),
    "odd_rows_7": PB((21, 18, 7), (1.0, 1.0, 1.0), np.deg2rad(np.linspace(0, 180, 36,
                                                                         endpoint=False) + 2),
                     centred(30, 1.0, 0.21), centred(7, 1.0)),
    "odd_rows_13": PB((40, 33, 13), (1.0, 1.0, 1.0), np.deg2rad(np.linspace(0, 180, 50,
                                                                           endpoint=False)),
                      centred(58, 1.0, -0.4), centred(13, 1.0)),
    "one_group": PB((512, 512, 16), (1.0, 1.0, 1.0), np.deg2rad(np.linspace(0, 180, 60,
                                                                           endpoint=False)),
                    centred(725, 1.0), centred(16, 1.0)),
}


def _geom(g):
    from ccpi.reconstruction.projectors import Geometry
    return Geometry(g.angles, g.h_pix, g.v_pix, g.origin, g.vsize, g.shape)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("name", list(CASES))
def test_tiled_bp_parity_adjoint_accumulate(name):
    from ccpi.reconstruction.projectors import backward_project, forward_project
    g = CASES[name]
    og = oracle.Geometry(g.angles, g.h_pix, g.v_pix, g.origin, g.vsize, g.shape)
    rng = np.random.default_rng(23)
    s = rng.random(g.sino_shape).astype(np.float32)
    x0 = rng.random(g.shape).astype(np.float32)
    print(f"\ntiled parallel BP {name} {g.shape}", flush=True)
    G = _geom(g)
    b = backward_project(G, torch.from_numpy(s).cuda()).cpu().numpy()
    bo = og.bp(s)
    assert _rel(b, bo) < 1e-5
    # accumulation into x0 (b2D adds to the volume)
    b2 = backward_project(G, torch.from_numpy(s).cuda(), torch.from_numpy(x0).cuda())
    assert _rel(b2.cpu().numpy(), og.bp(s, x0.copy())) < 1e-5
    again = backward_project(G, torch.from_numpy(s).cuda()).cpu().numpy()
    assert np.array_equal(again, b)
    y = forward_project(G, torch.from_numpy(x0).cuda()).cpu().numpy()
    lhs = float(np.vdot(y.astype(np.float64), s))
    rhs = float(np.vdot(x0.astype(np.float64), b))
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)
