"""Pins of the instrument set-up rules that both sides of every GPU
comparison share (CPU test).

The Python path derives its geometry with the reference's rules
(instruments.py restates them); a GPU test compares the device result with
the oracle on that same geometry, so a mistake in these rules would pass
there.  Here each rule is checked against what it is meant to achieve,
written independently of its restatement:

  * calc_v_alignment (instruments.cpp:35-63): the padded row count is the
    smallest multiple of 4 * ppv (8 * ppv for the cone, whose halves must be
    aligned) that holds the data rows, found by search, and the data sit in
    the middle (offset (padded - n) // 2);
  * the Diamond volume (diamond.cpp:479-501): the voxel square strictly
    contains every detector column, and one voxel less in x would not (or
    the size is the pixel count rounded up to whole voxels);
  * the XTek volume (xtek.cpp:760-798): the ray from the source to the top
    detector row passes the volume's far face below its top (the volume is
    tall enough for the whole cone); when the rule enlarges the voxels it
    solves that condition with equality before its (1 + 1 / nv) margin;
  * the scan ingest (diamond.cpp:394-451, xtek.cpp:519-562): the padded
    sinogram holds -log of the transposed data at the offset rows and zeros
    elsewhere, checked with numpy for a row count that needs padding on
    both sides.
"""

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")


def _padded_rows_by_search(n, ppv, cone):
    step = 4 * ppv * (2 if cone else 1)
    m = n
    while m % step:
        m += 1
    return m


@pytest.mark.parametrize("cone", [False, True])
@pytest.mark.parametrize("ppv", [1, 2, 3, 4])
def test_calc_v_alignment_is_the_smallest_aligned_padding(ppv, cone):
    from ccpi.reconstruction.instruments import calc_v_alignment
    for n in range(1, 130):
        npix, off = calc_v_alignment(n, ppv, cone)
        assert npix == _padded_rows_by_search(n, ppv, cone), (n, ppv, cone)
        assert 0 <= off and npix - n - off in (off, off + 1)


@pytest.mark.parametrize("nh,centre,ppv", [(160, 0.0, 1), (161, 0.0, 1), (160, 90.25, 1),
                                           (255, 0.0, 2), (300, 140.5, 3), (33, 0.0, 4)])
def test_diamond_volume_contains_the_detector(nh, centre, ppv):
    from ccpi.reconstruction import instruments as ins
    angles = np.linspace(0, 180, 10, endpoint=False)
    s = ins.diamond_setup((10, 8, nh), angles, centre, ppv)
    g = s.geometry
    nx, ny, _ = g.shape
    h = g.h_pixels
    half = g.voxel_size[0] * nx / 2.0
    assert nx == ny
    assert g.voxel_size[0] == pytest.approx((h[1] - h[0]) * ppv, rel=1e-15)
    # strictly contains every column's centre
    assert -half < h[0] and h[-1] < half
    # minimal: one voxel less fails, unless the size is the columns' span in
    # whole voxels (the rule's starting point)
    start = math.ceil(math.ceil(2 * max(abs(h[0]), abs(h[-1])) / (h[1] - h[0])) / ppv)
    smaller = g.voxel_size[0] * (nx - 1) / 2.0
    assert nx == start or not (-smaller < h[0] and h[-1] < smaller)


@pytest.mark.parametrize("shape,pix,sx,dx,mask", [
    ((720, 768, 768), 1.0, -500.0, 500.0, 128.0),       # corners near the source
    ((360, 384, 384), 2.0, -500.0, 500.0, 128.0),
    ((250, 512, 512), 0.390625, -250.0, 737.0, 25.15),  # mask sets the voxels
    ((100, 200, 150), 0.5, -300.0, 400.0, 60.0),
])
def test_xtek_volume_holds_the_cone(shape, pix, sx, dx, mask):
    from ccpi.reconstruction import instruments as ins
    angles = np.linspace(0, 360, shape[0], endpoint=False)
    s = ins.xtek_setup(shape, angles, 1, sx, dx, pix, pix, mask)
    g = s.geometry
    n, _, nz = g.shape
    vs = g.voxel_size[0]
    v_top = g.v_pixels[-1]
    far = vs * n / 2.0
    top = vs * nz / 2.0
    z_at_far_face = v_top * (far - sx) / (dx - sx)
    assert z_at_far_face <= top * (1 + 1e-12)
    mask_vs = 2.0 * mask / n
    if vs != mask_vs:
        # enlarged: without the (1 + 1/nv) margin the condition is an equality
        nv = len(g.v_pixels)
        v0 = vs / (1.0 + 1.0 / nv)
        assert v_top * (v0 * n / 2.0 - sx) / (dx - sx) == pytest.approx(v0 * nz / 2.0, rel=1e-12)
        assert v0 > mask_vs
    else:
        assert vs == pytest.approx(mask_vs, rel=1e-15)


def test_ingest_pads_rows_on_both_sides():
    from ccpi.reconstruction import instruments as ins
    rng = np.random.default_rng(3)
    na, nv_data, nh = 5, 6, 9          # 6 rows -> 8 padded rows, offset 1
    angles = np.linspace(0, 180, na, endpoint=False)
    scan = rng.uniform(0.2, 1.0, (na, nv_data, nh)).astype(np.float32)
    s = ins.diamond_setup(scan.shape, angles, 0.0, 1)
    assert s.geometry.sino_shape == (na, nh, 8) and s.data_v_offset == 1
    out = ins.diamond_read_scans(torch.from_numpy(scan), s, False).numpy()
    want = np.zeros((na, nh, 8), np.float32)
    want[:, :, 1:7] = -np.log(scan).transpose(0, 2, 1)
    assert np.allclose(out, want, rtol=1e-6, atol=0)
    assert np.all(out[:, :, 0] == 0) and np.all(out[:, :, 7] == 0)
    logged = ins.diamond_read_scans(torch.from_numpy(scan), s, True).numpy()
    assert np.array_equal(logged[:, :, 1:7], scan.transpose(0, 2, 1))
    # XTek scans come as [angle][h][v]
    xs = rng.uniform(0.2, 1.0, (na, nh, nv_data)).astype(np.float32)
    sx = ins.xtek_setup(xs.shape, np.linspace(0, 360, na, endpoint=False), 1, -100.0, 100.0,
                        0.5, 0.5, 2.0)
    assert sx.data_v_offset == 1 and sx.geometry.sino_shape == (na, nh, 8)
    xo = ins.xtek_read_scans(torch.from_numpy(xs), sx, False).numpy()
    assert np.allclose(xo[:, :, 1:7], -np.log(xs), rtol=1e-6, atol=0)
    assert np.all(xo[:, :, [0, 7]] == 0)
