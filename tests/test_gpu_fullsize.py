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
