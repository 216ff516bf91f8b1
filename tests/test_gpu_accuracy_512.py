"""The headline workload's accuracy against the oracle runs (GPU test).

Cone-beam CGLS on the 512^3 scan (0.5 mm voxels, 720 projections, 768^2
detector of 1 mm, SOD 500, SDD 1000; 30-ellipsoid phantom, seed 2, Poisson
noise I0 = 1e4): the oracle's FP64 and float (the reference's arithmetic)
CGLS for 10 iterations on the same data are committed in
tests/golden/accuracy_cone512.npz (tests/golden/make_accuracy.py).  The
device builds the same data (SHA-256 of the noise-free and noisy sinograms
must match) and runs 10 iterations with FP64 reductions: per iteration its
RMSE against the phantom equals the FP64 oracle's to 1e-6 and ||A^T r|| to
1e-5, and it is no further from the FP64 oracle's RMSE than the float
oracle's is, plus 1e-7 relative.
"""

import hashlib
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_cone512_cgls_matches_the_fp64_oracle():
    from ccpi.reconstruction import workloads as W
    from ccpi.reconstruction.projectors import CglsSession, forward_project
    path = os.path.join(GOLDEN, "accuracy_cone512.npz")
    f = np.load(path)
    g = W.cone_geometry(**W.CONE_512)
    print("\n512^3 data", flush=True)
    ph = torch.from_numpy(W.ellipsoid_phantom_host(g, 30, seed=2, dens=(0.0, 0.05))).cuda()
    clean = forward_project(g, ph).cpu().numpy()
    assert hashlib.sha256(clean.tobytes()).hexdigest() == str(f["sha256_clean"])
    bn = W.poisson_log_host(clean, 1e4, seed=2)
    del clean
    assert hashlib.sha256(bn.tobytes()).hexdigest() == str(f["sha256_noisy"])
    b = torch.from_numpy(bn).cuda()
    iters = len(f["rmse_f64"])
    x = torch.zeros(g.shape, dtype=torch.float32, device="cuda")
    s = CglsSession(g, b, x, iters)
    rmse = []
    for _ in range(iters):
        s.iterate(1)
        rmse.append(float(torch.sqrt(torch.mean((x - ph).double() ** 2))))
    norms = np.asarray(s.norms())
    s.close()
    r64, r32, n64 = f["rmse_f64"], f["rmse_f32"], f["norms_f64"]
    for it in range(iters):
        print(f"it {it + 1}: rmse device {rmse[it]:.9e} fp64 oracle {r64[it]:.9e} "
              f"float oracle {r32[it]:.9e}", flush=True)
    rmse = np.array(rmse)
    assert np.all(np.abs(rmse / r64 - 1) < 1e-6)
    assert np.all(np.abs(norms / n64 - 1) < 1e-5)
    assert np.all(np.abs(rmse - r64) <= np.abs(r32 - r64) + 1e-7 * r64)
