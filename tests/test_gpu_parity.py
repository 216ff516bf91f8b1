"""HIP projectors and iteration bodies against the CPU oracle (GPU tests).

Tolerances (float32 path, relative to the output's max magnitude):
  * cone FP: weights are bit-identical to the model and the per-row
    summation order is the reference's, so the result agrees to 1e-6;
  * parallel FP: whole-grid walk vs the reference's 32x32 tiles changes
    the summation order: 1e-5;
  * BPs: LDS / register accumulation order: 1e-5;
  * solvers (5 iterations): FP64 reductions vs the reference's float
    ones: 1e-4 for CGLS/SIRT/MLEM iterates.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle
from geometries import ALL_CB, ALL_PB, oracle_geom, ellipsoid_phantom

pytestmark = pytest.mark.gpu

ALL = ALL_PB | ALL_CB


def _geom(g):
    from ccpi.reconstruction.projectors import Geometry
    return Geometry(g.angles, g.h_pix, g.v_pix, g.origin, g.vsize, g.shape, cone=g.cone,
                    source=g.source, detector_x=g.det_x)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


@pytest.mark.parametrize("name", list(ALL))
def test_fp_parity(name, dev):
    from ccpi.reconstruction.projectors import forward_project
    g = ALL[name]()
    rng = np.random.default_rng(21)
    x = rng.random(g.shape).astype(np.float32)
    y0 = rng.random(g.sino_shape).astype(np.float32)  # FP accumulates
    y = forward_project(_geom(g), torch.from_numpy(x).to(dev), torch.from_numpy(y0).to(dev))
    yo = oracle_geom(g).fp(x, y0.copy())
    tol = 1e-6 if g.cone else 1e-5
    assert _rel(y.cpu().numpy(), yo) < tol


@pytest.mark.parametrize("name", list(ALL))
def test_bp_parity(name, dev):
    from ccpi.reconstruction.projectors import backward_project
    g = ALL[name]()
    rng = np.random.default_rng(22)
    s = rng.random(g.sino_shape).astype(np.float32)
    x0 = rng.random(g.shape).astype(np.float32)  # BP accumulates
    st = torch.from_numpy(s).to(dev)
    x = backward_project(_geom(g), st, torch.from_numpy(x0).to(dev))
    xo = oracle_geom(g).bp(s, x0.copy())
    assert _rel(x.cpu().numpy(), xo) < 1e-5
    assert torch.equal(st.cpu(), torch.from_numpy(s)), "BP must not modify the sinogram"


@pytest.mark.parametrize("name", list(ALL))
def test_adjoint_identity_gpu(name, dev):
    from ccpi.reconstruction.projectors import backward_project, forward_project
    g = ALL[name]()
    gg = _geom(g)
    rng = np.random.default_rng(23)
    x = torch.from_numpy(rng.standard_normal(g.shape).astype(np.float32)).to(dev)
    y = torch.from_numpy(rng.standard_normal(g.sino_shape).astype(np.float32)).to(dev)
    ax = forward_project(gg, x).double()
    aty = backward_project(gg, y).double()
    lhs = float((ax * y.double()).sum())
    rhs = float((x.double() * aty).sum())
    scale = float((forward_project(gg, x.abs()).double() * y.double().abs()).sum())
    assert abs(lhs - rhs) < 2e-5 * scale


def test_bp_deterministic(dev):
    from ccpi.reconstruction.projectors import backward_project
    g = ALL_CB["cb_small"]()
    rng = np.random.default_rng(24)
    s = torch.from_numpy(rng.random(g.sino_shape).astype(np.float32)).to(dev)
    a = backward_project(_geom(g), s)
    b = backward_project(_geom(g), s)
    assert torch.equal(a, b)


@pytest.mark.parametrize("name", ["pb_small", "cb_small", "cb_clip"])
def test_cgls_parity(name, dev):
    from ccpi.reconstruction.projectors import cgls
    g = ALL[name]()
    og = oracle_geom(g)
    x_true = ellipsoid_phantom(g.shape, g.vsize, g.origin, 4, seed=3)
    b = og.fp(x_true)
    xo, ro, no = oracle.cgls(og, b, 5)
    x, r, n = cgls(_geom(g), torch.from_numpy(b).to(dev), 5)
    assert _rel(x.cpu().numpy(), xo) < 1e-4
    assert _rel(r.cpu().numpy(), ro) < 1e-4
    # residual norms against the FP64 oracle: the float oracle's partial
    # sums are combined with OpenMP atomics (blas.hpp), so their low bits
    # depend on the thread count
    _, _, no64 = oracle.cgls(og, b.astype(np.float64), 5, dtype=np.float64)
    assert np.allclose(n, no64, rtol=2e-4)
    # reference-precision reductions give the same iterates
    x32, _, n32 = cgls(_geom(g), torch.from_numpy(b).to(dev), 5, reduce_fp32=True)
    assert _rel(x32.cpu().numpy(), xo) < 1e-4


@pytest.mark.parametrize("name", ["pb_small", "cb_small"])
def test_sirt_mlem_parity(name, dev):
    from ccpi.reconstruction.projectors import mlem, sirt
    g = ALL[name]()
    og = oracle_geom(g)
    x_true = ellipsoid_phantom(g.shape, g.vsize, g.origin, 4, seed=4)
    b = og.fp(x_true)
    xs = sirt(_geom(g), torch.from_numpy(b).to(dev), 5)
    assert _rel(xs.cpu().numpy(), oracle.sirt(og, b, 5)) < 1e-4
    xm = mlem(_geom(g), torch.from_numpy(b).to(dev), 5)
    assert _rel(xm.cpu().numpy(), oracle.mlem(og, b, 5)) < 1e-4
