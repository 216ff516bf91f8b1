"""MLEM with the projections' sums accumulated in double (CCPI_ACCUM_FP64)
against the FP64 oracle (GPU test).

The option keeps the float projector weights and element-wise steps and
accumulates A x, A^T (b / A x) and A^T 1 in double, rounding once.  On a
cone case (96^3, 120 angles, 144^2 detector, the 512^3 scan's ratios) and a
parallel case (the tiled backprojector's angle groups included), after 10
iterations its result is compared with the FP64 oracle's MLEM: it must be
no further from it than the default float accumulation is (printed with
-s), and both stay within the float oracle's own distance to the FP64
oracle plus 1e-5.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle

pytestmark = pytest.mark.gpu


def _cases():
    from ccpi.reconstruction import workloads as W
    cone = W.cone_geometry(96, 256.0 / 96, 120, 144, 768.0 / 144, 500.0, 1000.0)
    par = W.parallel_geometry(128, 16, 180, 183)
    return {"cone96": cone, "parallel128": par}


@pytest.mark.parametrize("name", ["cone96", "parallel128"])
def test_mlem_fp64_accumulation_against_fp64_oracle(name):
    from ccpi.reconstruction import workloads as W
    from ccpi.reconstruction.projectors import forward_project, mlem
    g = _cases()[name]
    og = oracle.Geometry(g.angles, g.h_pixels, g.v_pixels, g.origin, g.voxel_size, g.shape,
                         cone=g.cone, source=g.source, det_x=g.detector_x)
    ph = W.ellipsoid_phantom_host(g, 12, seed=8, dens=(0.0, 0.05))
    clean = forward_project(g, torch.from_numpy(ph).cuda()).cpu().numpy()
    bn = W.poisson_log_host(clean, 1e4, seed=8)
    b = torch.from_numpy(bn).cuda()
    iters = 10
    x64 = oracle.mlem(og, bn.astype(np.float64), iters, dtype=np.float64)
    x32 = oracle.mlem(og, bn, iters).astype(np.float64)
    xd = mlem(g, b, iters).cpu().numpy().astype(np.float64)
    xa = mlem(g, b, iters, accum_fp64=True).cpu().numpy().astype(np.float64)
    nrm = np.linalg.norm(x64)
    e_ref = np.linalg.norm(x32 - x64) / nrm
    e_def = np.linalg.norm(xd - x64) / nrm
    e_acc = np.linalg.norm(xa - x64) / nrm
    print(f"\n{name}: distance to FP64 MLEM: float oracle {e_ref:.3e}, device float sums "
          f"{e_def:.3e}, device FP64 sums {e_acc:.3e}", flush=True)
    assert e_acc <= e_def * 1.05 + 1e-7
    assert e_def <= e_ref + 1e-5
    assert e_acc <= e_ref + 1e-5
