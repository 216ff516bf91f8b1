"""Pins of the oracle's XTek centre-of-rotation search
(Core/src/Readers/xtek.cpp:566-758).

A cone-beam scan measured with the whole set-up translated by delta in y
(source at y = delta, every h pixel + delta) and searched with the nominal
geometry must return y_centre = delta, the shift the reference then adds
back (xtek.cpp:755-756).  Recovered within 0.005 mm (< 1 % of a pixel) for
shifts of either sign; a wrong sign, a dropped 2 gamma_c term or a wrong
conjugate-angle wrap fails by the shift itself.  A scan whose conjugate rays
never land on the detector reports the reference's failure.
"""

import numpy as np
import pytest

import oracle
from geometries import ellipsoid_phantom


def _scan(delta, na=90, nh=48, nv=16):
    from ccpi.reconstruction import instruments as ins
    ang = np.linspace(0, 360, na, endpoint=False).astype(np.float32)
    setup = ins.xtek_setup((na, nh, nv), ang, 1, -60.0, 60.0, 0.5, 1.0, 5.5)
    g = setup.geometry
    ph = ellipsoid_phantom(g.shape, g.voxel_size, g.origin, 5, seed=9, dens=(0.01, 0.05))
    true = oracle.Geometry(g.angles, g.h_pixels + delta, g.v_pixels, g.origin, g.voxel_size,
                           g.shape, cone=True, source=(g.source[0], delta, 0.0),
                           det_x=g.detector_x)
    nominal = oracle.Geometry(g.angles, g.h_pixels, g.v_pixels, g.origin, g.voxel_size,
                              g.shape, cone=True, source=g.source, det_x=g.detector_x)
    return nominal, true.fp(ph).astype(np.float32), nv // 2


@pytest.mark.parametrize("delta", [0.0, 0.13, -0.37, 0.8])
def test_find_centre_recovers_a_translated_setup(delta):
    og, y, vs = _scan(delta)
    yc, M = oracle.find_centre(og, y, vs)
    assert abs(yc - delta) < 5e-3, (yc, delta)
    # the last level's 21 candidates straddle the minimum
    assert 0 < int(np.argmin(M[4])) < 20


def test_find_centre_failure_without_conjugate_rays():
    og, y, vs = _scan(0.0, na=4, nh=2)
    og.h_pix = np.array([-400.0, -399.0])
    yc, _ = oracle.find_centre(og, y, vs)
    assert yc is None
