"""Size-independent properties at the full 512^3 cone-beam workload size
(720 projections over 360 deg, 768^2 detector of 1 mm, SOD 500, SDD 1000,
0.5 mm voxels), where the oracle would take minutes (GPU test):

  * A 1 is the chord length of every ray through the volume box: checked
    against the closed-form ray-box lengths (tests/bruteforce.py) on four
    whole projections, within 5e-6 of the maximum chord (2.2e-7 measured;
    float sums of ~10^3 exact segment lengths per ray);
  * the adjoint identity <A x, y> = <x, A^T y> for random x, y (FP64 dot
    products), within 1e-5;
  * exact scaling: A(2x) = 2 A(x) and A^T(2y) = 2 A^T y bit for bit (a power
    of two commutes with every float rounding);
  * the backprojection is deterministic run to run.

The same properties hold for the BP's phased form (more than 4 detector rows
per slice) on a wide cone at size: a 360 x 384^2 scan of 2 mm pixels (SOD
500, SDD 1000) through the XTek volume rule (instruments.xtek_setup,
xtek.cpp:760-798) gives 384^3 voxels of 1.62 mm whose corners come within
60 mm of the source, 14.5 rows per slice at most: S = 8 in two phases.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import bruteforce as bf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cone512():
    from ccpi.reconstruction import workloads as W
    return W.cone_geometry(**W.CONE_512)


def test_fp_of_ones_is_the_box_chord(cone512):
    from ccpi.reconstruction.projectors import forward_project
    g = cone512
    print("\n512^3 FP of ones", flush=True)
    y = forward_project(g, torch.ones(g.shape, device="cuda"))
    lo = tuple(g.origin)
    hi = tuple(g.origin[q] + g.shape[q] * g.voxel_size[q] for q in range(3))
    worst = 0.0
    for a in (0, 97, 360, 611):
        want = bf.cone_box_lengths(g.source, g.detector_x, g.angles[a], g.h_pixels,
                                   g.v_pixels, [(lo, hi)])
        got = y[a].double().cpu().numpy()
        worst = max(worst, float(np.abs(got - want).max() / want.max()))
    print(f"max deviation from the chord {worst:.2e} of the longest chord", flush=True)
    assert worst < 5e-6


def test_adjoint_scaling_and_determinism(cone512):
    from ccpi.reconstruction.projectors import backward_project, forward_project
    g = cone512
    gen = torch.Generator(device="cuda").manual_seed(12)
    x = torch.rand(g.shape, device="cuda", generator=gen)
    s = torch.rand(g.sino_shape, device="cuda", generator=gen)
    print("\n512^3 adjoint", flush=True)
    y = forward_project(g, x)
    b = backward_project(g, s)
    lhs = float(torch.sum(y.double() * s.double()))
    rhs = float(torch.sum(x.double() * b.double()))
    print(f"<Ax,y> {lhs:.9e} <x,A^T y> {rhs:.9e}", flush=True)
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)
    assert torch.equal(forward_project(g, 2 * x), 2 * y)
    del y
    assert torch.equal(backward_project(g, 2 * s), 2 * b)
    assert torch.equal(backward_project(g, s), b)


def test_wide_cone_phased_bp_adjoint_scaling_and_determinism():
    from ccpi.reconstruction import instruments as ins
    from ccpi.reconstruction.projectors import backward_project, forward_project
    ang = np.linspace(0, 360, 360, endpoint=False).astype(np.float32)
    g = ins.xtek_setup((360, 384, 384), ang, 1, -500.0, 500.0, 2.0, 2.0, 128.0).geometry
    assert g.shape == (384, 384, 384)
    gen = torch.Generator(device="cuda").manual_seed(13)
    x = torch.rand(g.shape, device="cuda", generator=gen)
    s = torch.rand(g.sino_shape, device="cuda", generator=gen)
    print("\nwide-cone 384^3 (phased BP) adjoint", flush=True)
    y = forward_project(g, x)
    b = backward_project(g, s)
    lhs = float(torch.sum(y.double() * s.double()))
    rhs = float(torch.sum(x.double() * b.double()))
    print(f"<Ax,y> {lhs:.9e} <x,A^T y> {rhs:.9e}", flush=True)
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)
    assert torch.equal(backward_project(g, 2 * s), 2 * b)
    assert torch.equal(backward_project(g, s), b)


def test_parallel_config1_size_parity_and_chords():
    """The plain parallel-beam workload (256 x 256 x 16, 180 angles over
    [0, 180 deg), 367 x 16 detector): FP and BP against the oracle (1e-5 of
    the maximum, the summation-order tolerance of test_gpu_parity) and A 1
    against the closed-form chord of each ray through the 256 mm square
    (rays along an edge of the square excepted)."""
    import oracle
    from ccpi.reconstruction import workloads as W
    from ccpi.reconstruction.projectors import backward_project, forward_project
    g = W.parallel_geometry(256, 16, 180, 367)
    og = oracle.Geometry(g.angles, g.h_pixels, g.v_pixels, g.origin, g.voxel_size, g.shape)
    rng = np.random.default_rng(13)
    x = rng.random(g.shape).astype(np.float32)
    s = rng.random(g.sino_shape).astype(np.float32)
    print("\nparallel 256x256x16 parity", flush=True)
    y = forward_project(g, torch.from_numpy(x).cuda()).cpu().numpy()
    yo = og.fp(x)
    assert np.abs(y - yo).max() < 1e-5 * np.abs(yo).max()
    b = backward_project(g, torch.from_numpy(s).cuda()).cpu().numpy()
    bo = og.bp(s)
    assert np.abs(b - bo).max() < 1e-5 * np.abs(bo).max()
    ones = forward_project(g, torch.ones(g.shape, device="cuda")).double().cpu().numpy()
    half = 128.0
    for a in (0, 45, 90, 131):
        c, sn = np.cos(g.angles[a]), np.sin(g.angles[a])
        # ray through (-sn h, c h) along (c, sn): slab intervals of the square
        h = g.h_pixels
        t = [bf._slab(-sn * h, c, -half, half), bf._slab(c * h, sn, -half, half)]
        chord = np.maximum(0.0, np.minimum(t[0][1], t[1][1]) - np.maximum(t[0][0], t[1][0]))
        chord[~np.isfinite(chord)] = 0.0
        # a ray running along an edge of the square is a tie of the model
        # (half of its length is credited to the outside row of cells)
        edge = ((abs(sn) < 1e-9) & (np.abs(np.abs(c * h) - half) < 1e-9)) | \
               ((abs(c) < 1e-9) & (np.abs(np.abs(sn * h) - half) < 1e-9))
        dev = np.abs(ones[a] - chord[:, None])[~edge]
        assert dev.max() < 1e-5 * chord.max()
