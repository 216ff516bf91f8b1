"""The Reconstructor front end (CPU tests): centre-of-rotation estimate on
parallel-beam scans with a known axis, and the parameter store.

Scans are projected by the oracle through the Diamond set-up with
rotation_centre rc (the axis rc pixels from the detector's left edge,
diamond.cpp:35-86); find_center_vo must return rc to within its fine step.
"""

import numpy as np
import pytest

import oracle
from geometries import ellipsoid_phantom


def _scan(rc, na=120, nh=96, nv=4):
    from ccpi.reconstruction import instruments as ins
    ang = np.linspace(0, 180, na, endpoint=False).astype(np.float32)
    setup = ins.diamond_setup((na, nv, nh), ang, rc, 1)
    g = setup.geometry
    og = oracle.Geometry(g.angles, g.h_pixels, g.v_pixels, g.origin, g.voxel_size, g.shape)
    ph = ellipsoid_phantom(g.shape, g.voxel_size, g.origin, 6, seed=11, dens=(0.2, 1.0))
    sino = og.fp(ph)
    return np.ascontiguousarray(sino.transpose(0, 2, 1)), ang


@pytest.mark.parametrize("rc", [48.0, 51.7, 44.2])
def test_find_center_vo_recovers_the_axis(rc):
    from ccpi.reconstruction.Reconstructor import find_center_vo
    tomo, _ = _scan(rc)
    got = find_center_vo(tomo, smin=-10, smax=10, srad=2, step=0.25)
    assert abs(got - rc) <= 0.25, (got, rc)


def test_parameter_store():
    from ccpi.reconstruction.Reconstructor import IterativeReconstructor, Reconstructor
    from ccpi.reconstruction.parallelbeam import alg
    r = Reconstructor.create(IterativeReconstructor, algorithm=alg.sirt, iterations=3)
    assert isinstance(r, IterativeReconstructor)
    assert r.getParameter(["algorithm", "iterations", "resolution"]) == [alg.sirt, 3, 1]
    with pytest.raises(KeyError):
        r.setParameter(bogus=1)
    with pytest.raises(KeyError):
        r.getParameter("angles")
    with pytest.raises(Exception, match="normalized projections"):
        r.iterate()
    assert "cgls_TVreg" in Reconstructor.printAvailableAlgorithms()
