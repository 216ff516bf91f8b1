"""XTek centre-of-rotation search on the device against the oracle (GPU
test).

The device search forms the same float terms and adds them in the
reference's sequential order, so the chosen centre is identical to the
oracle's, bit for bit: on translated cone scans (both shift signs), with
the angles in scan order (binary search) and permuted (the reference's
linear scan), and on a 180 x 256 detector row of noise.  The 720 x 768
size of the 512^3 workload is timed (printed with -s).
"""

import time

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle
from test_oracle_centre_pins import _scan

pytestmark = pytest.mark.gpu


def _device_centre(og, y, v_slice):
    from ccpi.reconstruction import instruments as ins
    from ccpi.reconstruction.projectors import Geometry
    g = Geometry(og.angles, og.h_pix, og.v_pix, og.origin, og.vsize, og.shape, cone=True,
                 source=og.source, detector_x=og.det_x)
    setup = ins.Setup(g, og.v_pix.size, 0, 0.0)
    yc = ins.find_centre(setup, torch.from_numpy(y).cuda(), v_slice)
    return yc, g


@pytest.mark.parametrize("delta", [0.0, 0.13, -0.37])
def test_find_centre_matches_oracle(delta):
    og, y, vs = _scan(delta)
    print(f"\nfind_centre delta {delta}", flush=True)
    want, _ = oracle.find_centre(og, y, vs)
    got, g = _device_centre(og, y, vs)
    assert got == want
    assert g.source[1] == og.source[1] + got
    assert np.array_equal(g.h_pixels, og.h_pix + got)


def test_find_centre_unsorted_angles():
    og, y, vs = _scan(0.21)
    perm = np.random.default_rng(2).permutation(og.angles.size)
    og.angles = og.angles[perm].copy()
    y = np.ascontiguousarray(y[perm])
    want, _ = oracle.find_centre(og, y, vs)
    got, _ = _device_centre(og, y, vs)
    assert got == want


def test_find_centre_noise_row_and_timing():
    from ccpi.reconstruction import instruments as ins
    na, nh, nv = 180, 256, 4
    ang = np.linspace(0, 360, na, endpoint=False).astype(np.float32)
    setup = ins.xtek_setup((na, nh, nv), ang, 1, -500.0, 500.0, 1.0, 1.0, 60.0)
    g = setup.geometry
    og = oracle.Geometry(g.angles, g.h_pixels, g.v_pixels, g.origin, g.voxel_size, g.shape,
                         cone=True, source=g.source, det_x=g.detector_x)
    y = np.random.default_rng(3).random(g.sino_shape).astype(np.float32)
    want, _ = oracle.find_centre(og, y, 1)
    got, _ = _device_centre(og, y, 1)
    assert got == want
    na, nh, nv = 720, 768, 4
    ang = np.linspace(0, 360, na, endpoint=False).astype(np.float32)
    setup = ins.xtek_setup((na, nh, nv), ang, 1, -500.0, 500.0, 1.0, 1.0, 128.0)
    y = torch.rand(setup.geometry.sino_shape, device="cuda",
                   generator=torch.Generator(device="cuda").manual_seed(4))
    ins.find_centre(setup, y, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ins.find_centre(setup, y, 1)
    print(f"\nfind_centre 720 x 768: {1e3 * (time.perf_counter() - t0):.1f} ms", flush=True)
