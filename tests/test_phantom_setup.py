"""CPU checks of the XTek phantom scan set-up and of the closed form its GPU
test uses.

The metadata values are the reference's only value assertion
(Wrappers/Python/test/test_conebeam.py:17-27): 250 projections of 512 x 512
pixels, source 250.0, detector 987.0, pixel 0.390625 mm, mask radius
25.151547886198443.  The closed-form ray-box lengths are pinned against the
voxel-by-voxel exact matrix on a small grid.
"""

import numpy as np

import bruteforce as bf
from geometries import centred


def test_phantom_metadata_matches_reference_assertions():
    from ccpi.reconstruction import conebeam
    g, phi, mask = conebeam.phantom_setup()
    assert g.sino_shape == (250, 512, 512)
    assert g.shape == (512, 512, 512)
    sx, dx, ph, pv, mr = conebeam._phantom_scalars(g, mask)
    assert (sx, dx, ph, pv) == (250.0, 987.0, 0.390625, 0.390625)
    assert abs(mr - 25.151547886198443) < 1e-12
    assert g.h_pixels[0] == -99.8046875 and g.h_pixels[-1] == 99.8046875
    assert abs(g.voxel_size[0] * 512 - 2 * mr) < 1e-12
    assert abs(g.origin[0] + mr) < 1e-12
    assert abs(phi[1] - 2 * np.pi / 250) < 1e-15


def test_box_lengths_match_voxel_sums():
    shape, vs = (6, 5, 4), (1.0, 1.2, 0.8)
    origin = (-3.0, -3.0, -1.6)
    h, v = centred(9, 0.9), centred(7, 0.7)
    src, det = (-20.0, 0.3, 0.1), 25.0
    box = ((1, 4), (2, 5), (0, 3))
    lo = tuple(origin[q] + box[q][0] * vs[q] for q in range(3))
    hi = tuple(origin[q] + box[q][1] * vs[q] for q in range(3))
    for ang in (0.0, 0.7, 2.9):
        A = bf.cone_matrix(src, det, [ang], h, v, origin, vs, shape)
        mask = np.zeros(shape)
        mask[box[0][0]:box[0][1], box[1][0]:box[1][1], box[2][0]:box[2][1]] = 1
        want = (A @ mask.ravel()).reshape(len(h), len(v))
        got = bf.cone_box_lengths(src, det, ang, h, v, [(lo, hi)])
        assert np.abs(got - want).max() < 1e-12
