"""Precision of the CGLS reductions against the FP64 oracle (GPU test).

Cone geometry with the ratios of the 768^3 / 1200-angle precision workload
scaled down to 80^3 / 150 angles / 120^2 pixels (same field of view, SOD
500 mm, SDD 1000 mm), 50-ellipsoid phantom, Poisson noise I0 = 5e3, 100
CGLS iterations.  Variants: FP32 projectors with FP64 reductions (default),
FP32 projectors with FP32 reductions, the oracle's float CGLS (the
reference's arithmetic: float weights and float partial sums, blas.hpp:
204-284, cgls.cpp:26-157) and the oracle's full-FP64 CGLS.  Each variant's
relative distance to the FP64 iterate is its error; printed with -s.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle

pytestmark = pytest.mark.gpu


def test_fp64_reductions_track_the_fp64_oracle():
    from ccpi.reconstruction import workloads as W
    from ccpi.reconstruction.projectors import CglsSession
    n, na, nd = 80, 150, 120
    g = W.cone_geometry(n, 256.0 / n, na, nd, 768.0 / nd, 500.0, 1000.0)
    og = oracle.Geometry(g.angles, g.h_pixels, g.v_pixels, g.origin, g.voxel_size, g.shape,
                         cone=True, source=g.source, det_x=g.detector_x)
    ph = W.ellipsoid_phantom(g, 50, seed=5, dens=(0.0, 0.05), device="cuda")
    ph_np = ph.cpu().numpy().astype(np.float64)
    clean = og.fp(ph_np)
    rng = np.random.default_rng(5)
    counts = rng.poisson(5e3 * np.exp(-clean))
    b = (-np.log(np.maximum(counts, 1) / 5e3)).astype(np.float32)
    iters = 100
    x64, _, n64 = oracle.cgls(og, b.astype(np.float64), iters, dtype=np.float64)
    # the reference's own arithmetic: float weights, float partial sums
    x32, _, n32 = oracle.cgls(og, b, iters, dtype=np.float32)

    def dist(xv):
        return float(np.linalg.norm(xv.astype(np.float64) - x64) / np.linalg.norm(x64))

    res = {"reference-float (oracle)": (dist(x32), float(np.abs(n32 / n64 - 1).max()),
                                        float(np.sqrt(np.mean((x32 - ph_np) ** 2))), n32)}
    for name, fp32 in (("fp64-reductions", False), ("fp32-reductions", True)):
        x = torch.zeros(g.shape, dtype=torch.float32, device="cuda")
        s = CglsSession(g, torch.from_numpy(b).cuda(), x, iters, reduce_fp32=fp32)
        s.iterate(iters)
        norms = s.norms()
        s.close()
        xv = x.cpu().numpy().astype(np.float64)
        res[name] = (dist(xv), float(np.abs(norms / n64 - 1).max()),
                     float(np.sqrt(np.mean((xv - ph_np) ** 2))), norms)
    rm64 = float(np.sqrt(np.mean((x64 - ph_np) ** 2)))
    print(f"\nfp64 oracle rmse {rm64:.6e}; ||A^T r|| it 1/10/100: "
          f"{n64[0]:.4e} {n64[9]:.4e} {n64[-1]:.4e}")
    for k, (e, nrm, rm, _) in res.items():
        print(f"{k}: rel. distance to fp64 iterate {e:.3e}, max rel. norm dev {nrm:.3e}, "
              f"rmse {rm:.6e}")
    e_ref = res["reference-float (oracle)"][0]
    e64, _, rm_gpu, norms64 = res["fp64-reductions"]
    # ||A^T b|| (iteration 0) depends on the backprojection alone
    assert abs(norms64[0] / n64[0] - 1) < 1e-4
    # no further from the FP64 solve than the reference's float arithmetic
    # (the FP32 projector weights dominate both distances), and the same
    # RMSE against the phantom to 2 %
    assert e64 <= 1.25 * e_ref + 1e-4, (e64, e_ref)
    assert abs(rm_gpu / rm64 - 1) < 2e-2
