"""conebeam.alg.create_phantom / create_phantom_volume on the device (GPU
test).

The projections are pinned to the closed-form lengths of each source ->
pixel segment inside the three cubes (tests/bruteforce.py
cone_box_lengths, itself pinned against the voxel-by-voxel matrix) at four
angles, within 2e-5 of the maximum (the float alphas of the projector
model, as in the oracle pins); the metadata are the reference's assertions
(Wrappers/Python/test/test_conebeam.py:17-27).
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import bruteforce as bf

pytestmark = pytest.mark.gpu


def _boxes(g, cubes):
    o, vs = g.origin, g.voxel_size
    return [tuple(tuple(o[q] + c[q][e] * vs[q] for q in range(3)) for e in (0, 1))
            for c in cubes]


def test_create_phantom_projections_are_cube_lengths():
    from ccpi.reconstruction import conebeam
    from ccpi.reconstruction.conebeam import alg
    print("\ncreate_phantom 512^3, 250 x 512^2", flush=True)
    pixels, angles, sx, dx, ph, pv, mr = alg.create_phantom()
    assert pixels.shape == (250, 512, 512) and pixels.dtype == np.float32
    assert angles.dtype == np.float32 and angles.shape == (250,)
    assert abs(angles[1] - 1.44) < 1e-5
    assert (sx, dx, ph, pv) == (250.0, 987.0, 0.390625, 0.390625)
    assert abs(mr - 25.151547886198443) < 1e-12
    g, phi, _ = conebeam.phantom_setup()
    boxes = _boxes(g, conebeam.PHANTOM_CUBES)
    for a in (0, 37, 125, 249):
        want = bf.cone_box_lengths(g.source, g.detector_x, phi[a], g.h_pixels, g.v_pixels,
                                   boxes)
        err = np.abs(pixels[a] - want).max()
        assert err < 2e-5 * want.max(), (a, err, want.max())


def test_create_phantom_volume():
    from ccpi.reconstruction.conebeam import alg
    vox, angles, h, v, sx, dx, ph, pv, mr = alg.create_phantom_volume()
    assert vox.shape == (512, 512, 512) and vox.dtype == np.float32
    assert vox.sum() == 3 * 82 ** 3
    assert vox[107, 107, 57] == 1 and vox[106, 107, 57] == 0 and vox[352, 352, 302] == 1
    assert vox[189, 188, 139] == 0 and vox[189, 189, 139] == 1
    assert h.dtype == np.float32 and h.shape == (512,)
    assert np.array_equal(h, v) and h[0] == np.float32(-99.8046875)
    assert (sx, dx) == (250.0, 987.0) and angles.shape == (250,)
