"""The Python drop-in modules (parallelbeam.alg / conebeam.alg) and the
differentiable projector against the oracle (GPU tests)."""

import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle
from geometries import ALL_CB, centred, oracle_geom

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _og(g):
    return oracle.Geometry(g.angles, g.h_pixels, g.v_pixels, g.origin, g.voxel_size, g.shape,
                           cone=g.cone, source=g.source, det_x=g.detector_x)


def test_parallel_cgls_on_reference_fixture():
    """The reference's smoke call, Wrappers/Python/test/test.py:8:
    cgls(sino, angles, 0.0, 1, 5, 1, False) on its own sino.npy/angles.npy,
    against the same ingest + oracle CGLS + final clamp."""
    from ccpi.reconstruction import instruments as ins
    from ccpi.reconstruction.parallelbeam import alg
    sino = np.load(os.path.join(GOLDEN, "sino.npy")).astype(np.float32)
    angles = np.load(os.path.join(GOLDEN, "angles.npy")).astype(np.float32)
    vox = alg.cgls(sino, angles, np.double(0.0), 1, 5, 1, False)
    setup = ins.diamond_setup(sino.shape, angles, 0.0, 1)
    g = setup.geometry
    b = np.zeros(g.sino_shape, np.float32)
    off = setup.data_v_offset
    b[:, :, off:off + sino.shape[1]] = -np.log(sino.transpose(0, 2, 1))
    xo, _, _ = oracle.cgls(_og(g), b, 5)
    xo = np.maximum(xo, 0)
    assert vox.shape == g.shape == (160, 160, 8)
    assert _rel(vox, xo) < 2e-4
    assert len(alg.cgls_norms()) == 5


def test_pb_forward_backward_project_layouts():
    from ccpi.reconstruction import instruments as ins
    from ccpi.reconstruction.parallelbeam import alg
    rng = np.random.default_rng(5)
    vol = rng.random((20, 20, 4)).astype(np.float32)
    ang = np.linspace(0, 180, 15, endpoint=False).astype(np.float32)
    y = alg.pb_forward_project(vol, ang, 1)
    setup = ins.diamond_setup((15, 4, 20), ang, 10.0, 1, shape=vol.shape)
    yo = _og(setup.geometry).fp(vol)
    assert y.shape == (15, 4, 20)
    assert _rel(y, yo.transpose(0, 2, 1)) < 1e-5
    x = alg.pb_backward_project(y, ang, 10.0, 1)
    setup2 = ins.diamond_setup(y.shape, ang, 10.0, 1)
    xo = _og(setup2.geometry).bp(np.ascontiguousarray(y.transpose(0, 2, 1)))
    assert _rel(x, xo) < 1e-5


def test_cone_create_sinogram():
    from ccpi.reconstruction.conebeam import alg
    g = ALL_CB["cb_small"]()
    rng = np.random.default_rng(6)
    vol = rng.random(g.shape).astype(np.float32)
    y = alg.create_sinogram(g.source[0], g.source[1], g.source[2], g.det_x, g.h_pix,
                            g.v_pix, g.angles, vol, g.origin, g.vsize)
    f = lambda a: np.asarray(a, np.float32).astype(np.float64)
    og = oracle.Geometry(f(g.angles), f(g.h_pix), f(g.v_pix), f(g.origin), f(g.vsize),
                         g.shape, cone=True, source=tuple(f(g.source)), det_x=float(f(g.det_x)))
    assert _rel(y, og.fp(vol)) < 1e-6


def test_cone_cgls_python_path_find_centre():
    """conebeam.alg.cgls on data measured with a centred source: the
    find_centre search lands within 0.01 mm of no shift and the result
    matches the oracle CGLS on the same (shifted) geometry."""
    from ccpi.reconstruction import instruments as ins
    from ccpi.reconstruction.conebeam import alg
    na, nh, nv = 36, 24, 16
    ang = np.linspace(0, 360, na, endpoint=False).astype(np.float32)
    setup = ins.xtek_setup((na, nh, nv), ang, 1, -60.0, 60.0, 1.0, 1.0, 5.5)
    g = setup.geometry
    from geometries import ellipsoid_phantom
    ph = ellipsoid_phantom(g.shape, g.voxel_size, g.origin, 5, seed=9, dens=(0.01, 0.05))
    y = _og(g).fp(ph).astype(np.float32)
    vox = alg.cgls(y, ang, np.zeros(1, np.float32), np.zeros(1, np.float32), 1, 60.0, 120.0,
                   1.0, 1.0, 5.5, False, np.zeros(3, np.float32), np.ones(3, np.float32), 4, 1,
                   True)
    g2 = alg.last_geometry()
    assert abs(g2.source[1]) < 0.01
    xo, _, _ = oracle.cgls(_og(g2), y, 4)
    assert _rel(vox, np.maximum(xo, 0)) < 2e-4


def test_autograd_gradient_matches_finite_differences():
    """d/dx 0.5||Ax - b||^2 through ForwardProjection (hand-written adjoint)
    against float64 central differences of the FP64 oracle on 64 voxels."""
    from ccpi.reconstruction.autograd import data_fidelity
    from ccpi.reconstruction.projectors import Geometry
    c = ALL_CB["cb_small"]()
    g = Geometry(c.angles, c.h_pix, c.v_pix, c.origin, c.vsize, c.shape, cone=True,
                 source=c.source, detector_x=c.det_x)
    og = oracle_geom(c)
    rng = np.random.default_rng(12)
    x0 = rng.random(c.shape)
    b = og.fp(rng.random(c.shape))
    xt = torch.tensor(x0, dtype=torch.float32, device="cuda", requires_grad=True)
    loss = data_fidelity(xt, g, torch.tensor(b, dtype=torch.float32, device="cuda"))
    loss.backward()
    grad = xt.grad.double().cpu().numpy()
    L = lambda x: 0.5 * np.sum((og.fp(x) - b) ** 2)
    eps = 1e-3
    idx = rng.choice(x0.size, 64, replace=False)
    for q in idx:
        e = np.zeros(x0.size)
        e[q] = eps
        e = e.reshape(x0.shape)
        fd = (L(x0 + e) - L(x0 - e)) / (2 * eps)
        assert abs(grad.ravel()[q] - fd) <= 1e-4 * max(1.0, abs(fd)) + 1e-3 * np.abs(grad).max()


def test_backprojection_function_backward_is_forward():
    from ccpi.reconstruction.autograd import backproject
    from ccpi.reconstruction.projectors import Geometry, forward_project
    c = ALL_CB["cb_clip"]()
    g = Geometry(c.angles, c.h_pix, c.v_pix, c.origin, c.vsize, c.shape, cone=True,
                 source=c.source, detector_x=c.det_x)
    rng = np.random.default_rng(13)
    y = torch.tensor(rng.random(c.sino_shape), dtype=torch.float32, device="cuda",
                     requires_grad=True)
    w = torch.tensor(rng.random(c.shape), dtype=torch.float32, device="cuda")
    (backproject(y, g) * w).sum().backward()
    assert torch.allclose(y.grad, forward_project(g, w), rtol=1e-5, atol=1e-6)


def test_iterative_reconstructor_front_end():
    """Reconstructor.create(IterativeReconstructor, ...) runs the chosen
    parallelbeam.alg solver with the given parameters (Reconstructor.py:
    114-160) and, without a centre, estimates one with find_center_vo."""
    from ccpi.reconstruction.parallelbeam import alg
    from ccpi.reconstruction.Reconstructor import IterativeReconstructor, Reconstructor
    sino = np.load(os.path.join(GOLDEN, "sino.npy")).astype(np.float32)
    angles = np.load(os.path.join(GOLDEN, "angles.npy")).astype(np.float32)
    r = Reconstructor.create(IterativeReconstructor, algorithm=alg.cgls_tikhonov,
                             normalized_projections=sino, angles=angles,
                             center_of_rotation=80.0, iterations=3,
                             regularization_parameter=0.05)
    vol = r.iterate()
    want = alg.cgls_tikhonov(sino, angles, 80.0, 1, 3, 1, 0.05, np.zeros(3), False)
    assert np.array_equal(vol, want)
    r = Reconstructor.create(IterativeReconstructor, algorithm=alg.sirt,
                             normalized_projections=sino, angles=angles, iterations=2)
    vol = r.iterate()
    cor = r.getParameter("center_of_rotation")
    assert 40.0 < cor < 120.0
    assert vol.shape[2] == 8 and np.isfinite(vol).all()


@pytest.mark.parametrize("algo", ["cgls", "sirt", "mlem"])
def test_parallel_resolution_2_on_reference_fixture(algo):
    """resolution (pixels per voxel) 2 on the reference's fixture: half as
    many voxels per axis, two detector rows per slice (map type 2,
    p2D.cpp:22-61), with the rotation centre of the data (80.0)."""
    from ccpi.reconstruction import instruments as ins
    from ccpi.reconstruction.parallelbeam import alg
    sino = np.load(os.path.join(GOLDEN, "sino.npy")).astype(np.float32)
    angles = np.load(os.path.join(GOLDEN, "angles.npy")).astype(np.float32)
    vox = getattr(alg, algo)(sino, angles, np.double(80.0), 2, 3, 1, False)
    setup = ins.diamond_setup(sino.shape, angles, 80.0, 2)
    g = setup.geometry
    b = np.zeros(g.sino_shape, np.float32)
    off = setup.data_v_offset
    b[:, :, off:off + sino.shape[1]] = -np.log(sino.transpose(0, 2, 1))
    og = _og(g)
    if algo == "cgls":
        xo, _, _ = oracle.cgls(og, b, 3)
    else:
        xo = getattr(oracle, algo)(og, b, 3)
    assert vox.shape == g.shape and g.shape[0] == 80 and g.shape[2] == 4
    assert _rel(vox, np.maximum(xo, 0)) < 2e-4


@pytest.mark.parametrize("ppv", [1, 2])
def test_parallel_cgls_with_padded_rows(ppv):
    """Six detector rows of the reference's fixture (rows 1..6): the set-up
    pads them to 8 rows with one zero row on each side (data offset 1, the
    smallest multiple of 4 * ppv rows that holds them), so the data rows sit
    at sinogram rows 1..6; the oracle gets that sinogram built here by hand."""
    from ccpi.reconstruction import instruments as ins
    from ccpi.reconstruction.parallelbeam import alg
    sino = np.load(os.path.join(GOLDEN, "sino.npy")).astype(np.float32)[:, 1:7, :]
    angles = np.load(os.path.join(GOLDEN, "angles.npy")).astype(np.float32)
    vox = alg.cgls(np.ascontiguousarray(sino), angles, np.double(80.0), ppv, 4, 1, False)
    g = ins.diamond_setup(sino.shape, angles, 80.0, ppv).geometry
    assert g.sino_shape[2] == 8
    b = np.zeros(g.sino_shape, np.float32)
    b[:, :, 1:7] = -np.log(sino.transpose(0, 2, 1))
    xo, _, _ = oracle.cgls(_og(g), b, 4)
    assert vox.shape == g.shape and g.shape[2] == 8 // ppv
    assert _rel(vox, np.maximum(xo, 0)) < 2e-4
