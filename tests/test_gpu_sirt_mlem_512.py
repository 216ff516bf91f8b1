"""Cone-beam SIRT and MLEM at the 512^3 workload size (GPU test): 512^3 of
0.5 mm, 720 projections over 360 deg, 768^2 detector of 1 mm, SOD 500,
SDD 1000, the 30-ellipsoid phantom (seed 2), its noise-free projection as
the data (b = A x_true >= 0), 50 iterations each, where the oracle would
take hours.  Properties the algorithms guarantee:

  * SIRT (sirt.cpp:15-80) is a projected gradient method in the norms of
    R = 1 / (A 1) and C = 1 / (A^T 1): from 0, the R-weighted residual
    ||b - A x||_R never increases (checked every 5 iterations, continuing
    from the returned volume), x stays >= 0 and its RMSE against the
    phantom falls;
  * MLEM (mlem.cpp:15-80) increases the Poisson likelihood every iteration
    for b >= 0: the Kullback-Leibler divergence KL(b || A x) after 10, 25
    and 50 iterations (each run from x = 1) never increases, x >= 0, and
    the RMSE falls.
Slack: 1e-6 relative for the float arithmetic of the bodies.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scan512():
    from ccpi.reconstruction import workloads as W
    from ccpi.reconstruction.projectors import forward_project
    g = W.cone_geometry(**W.CONE_512)
    ph = torch.from_numpy(W.ellipsoid_phantom_host(g, 30, seed=2, dens=(0.0, 0.05))).cuda()
    b = forward_project(g, ph)
    return g, ph, b


def _rmse(x, ph):
    return float(torch.sqrt(torch.mean((x - ph).double() ** 2)))


def test_sirt_512_weighted_residual_decreases(scan512):
    from ccpi.reconstruction.projectors import forward_project, sirt
    g, ph, b = scan512
    print("\n512^3 SIRT", flush=True)
    R = forward_project(g, torch.ones(g.shape, device="cuda")).double()
    R = torch.where(R > 0, 1.0 / R, torch.zeros_like(R))
    x = torch.zeros(g.shape, dtype=torch.float32, device="cuda")
    res, rm = [], []
    for _ in range(10):
        x = sirt(g, b, 5, x=x)
        r = b.double() - forward_project(g, x).double()
        res.append(float(torch.sqrt(torch.sum(R * r * r))))
        rm.append(_rmse(x, ph))
        print(f"||b - Ax||_R {res[-1]:.6e} rmse {rm[-1]:.6e}", flush=True)
    assert float(x.min()) >= 0.0
    assert all(res[q + 1] <= res[q] * (1 + 1e-6) for q in range(len(res) - 1))
    assert rm[-1] < rm[0] < _rmse(torch.zeros_like(ph), ph)


def test_mlem_512_kl_decreases(scan512):
    from ccpi.reconstruction.projectors import forward_project, mlem
    g, ph, b = scan512
    print("\n512^3 MLEM", flush=True)
    bd = b.double()
    kl, rm = [], []
    for n in (10, 25, 50):
        x = mlem(g, b, n)
        ax = forward_project(g, x).double()
        pos = bd > 0
        t = torch.where(pos, bd * torch.log(bd / torch.clamp(ax, min=1e-30)), torch.zeros_like(bd))
        kl.append(float(torch.sum(t - bd + ax)))
        rm.append(_rmse(x, ph))
        assert float(x.min()) >= 0.0
        print(f"{n} iterations: KL(b || Ax) {kl[-1]:.6e} rmse {rm[-1]:.6e}", flush=True)
    assert kl[1] <= kl[0] * (1 + 1e-6) and kl[2] <= kl[1] * (1 + 1e-6)
    assert rm[2] < rm[0]
