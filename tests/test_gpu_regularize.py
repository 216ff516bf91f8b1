"""Regularisation terms and regularised CGLS on the device against the
oracle (GPU tests).

  * Tikhonov / TV terms: bit-exact against the float oracle (same float
    operations in the same order; quotients and square roots correctly
    rounded on both sides), on shapes with 2-wide extents, planes that are
    not a multiple of the 256-thread block and more planes than one launch
    row covers;
  * regularised CGLS (5 iterations): Tikhonov iterates within 1e-4 of the
    float oracle relative to the output's max (the backprojection's
    summation order differs, as for plain CGLS), norms within 2e-4 of the
    FP64 oracle; TV iterates no further from the FP64 oracle than 1.5x the
    float oracle's own distance to it (TV's 1/sqrt(FLT_EPSILON) gain on
    flat regions makes float runs of any summation order part that far);
  * the Python drop-ins cgls_tikhonov / cgls_TVreg / cgls_conv (and their
    _step forms) on the reference's own sino.npy / angles.npy, and the
    norm_r handling of diamond_wrapper.cpp:95-104.
"""

import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle
from geometries import ALL_CB, ALL_PB, ellipsoid_phantom, oracle_geom

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ALL = ALL_PB | ALL_CB


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _geom(g):
    from ccpi.reconstruction.projectors import Geometry
    return Geometry(g.angles, g.h_pix, g.v_pix, g.origin, g.vsize, g.shape, cone=g.cone,
                    source=g.source, detector_x=g.det_x)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


SHAPES = [(2, 2, 2), (5, 7, 3), (9, 40, 17), (3, 2, 300), (5000, 2, 3), (33, 65, 64)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind", ["tikhonov", "tv"])
def test_terms_bit_exact(shape, kind, dev):
    from ccpi.reconstruction.projectors import TIKHONOV, TV, regularize
    k = TIKHONOV if kind == "tikhonov" else TV
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape).astype(np.float32)
    # a few flat runs so that min-mod sees zero and opposite-sign differences
    a[:, :, : shape[2] // 2] = np.round(a[:, :, : shape[2] // 2])
    print(f"\n{kind} shape {shape}", flush=True)
    got = regularize(k, torch.from_numpy(a).to(dev)).cpu().numpy()
    want = oracle.regularize(k, a)
    assert np.array_equal(got, want), np.abs(got - want).max()


def test_terms_reject_thin_volumes(dev):
    from ccpi.reconstruction._lib import CcpiError
    from ccpi.reconstruction.projectors import TV, regularize
    with pytest.raises(CcpiError):
        regularize(TV, torch.zeros((4, 4, 1), device=dev))


@pytest.mark.parametrize("name", ["pb_small", "cb_small", "cb_clip"])
@pytest.mark.parametrize("kind,lam", [("tikhonov", 0.2), ("tv", 1e-3), ("tv-unstable", 1e-2)])
def test_cgls_reg_parity(name, kind, lam, dev):
    from ccpi.reconstruction.projectors import TIKHONOV, TV, cgls_reg
    k = TIKHONOV if kind == "tikhonov" else TV
    g = ALL[name]()
    og = oracle_geom(g)
    x_true = ellipsoid_phantom(g.shape, g.vsize, g.origin, 4, seed=3)
    b = og.fp(x_true)
    print(f"\ncgls_reg {kind} {name}", flush=True)
    xo, ro, _ = oracle.cgls_reg(og, b, 5, k, lam)
    x64, _, no64 = oracle.cgls_reg(og, b.astype(np.float64), 5, k, lam, dtype=np.float64)
    x, r, n = cgls_reg(_geom(g), torch.from_numpy(b).to(dev), 5, k, lam)
    x32, _, _ = cgls_reg(_geom(g), torch.from_numpy(b).to(dev), 5, k, lam, reduce_fp32=True)
    x, r, x32 = x.cpu().numpy(), r.cpu().numpy(), x32.cpu().numpy()
    if kind != "tv-unstable":
        assert _rel(x, xo) < 1e-4
        assert _rel(r, ro) < 1e-4
        assert _rel(x32, xo) < 1e-4
        assert np.allclose(n, no64, rtol=2e-4)
    else:
        # x += lambda TV(x) is an explicit step whose gain on flat regions
        # is ~lambda * 8 / sqrt(FLT_EPSILON) (~2e2 here): rounding noise
        # grows each iteration, and float runs that sum in different orders
        # part by up to the float oracle's own distance to the FP64 oracle
        # (1e-3 on cb_clip; the terms themselves are bit-exact, see above).
        # The device result must be no further from the FP64 solve than the
        # reference-precision oracle is.
        e_ref = _rel(xo, x64)
        assert _rel(x, x64) <= 1.5 * e_ref + 1e-5, (_rel(x, x64), e_ref)
        assert _rel(x32, x64) <= 1.5 * e_ref + 1e-5, (_rel(x32, x64), e_ref)
        assert _rel(x, xo) <= 2.0 * e_ref + 1e-5
        assert np.allclose(n, no64, rtol=2e-3)


def test_cgls_reg_zero_lambda_is_cgls(dev):
    from ccpi.reconstruction.projectors import TV, cgls, cgls_reg
    g = ALL["cb_small"]()
    og = oracle_geom(g)
    b = torch.from_numpy(og.fp(ellipsoid_phantom(g.shape, g.vsize, g.origin, 4, seed=5))).to(dev)
    x0, r0, n0 = cgls(_geom(g), b, 4)
    x1, r1, n1 = cgls_reg(_geom(g), b, 4, TV, 0.0)
    assert torch.equal(x0, x1) and torch.equal(r0, r1)
    assert np.array_equal(n0, n1)


def _fixture():
    sino = np.load(os.path.join(GOLDEN, "sino.npy")).astype(np.float32)
    angles = np.load(os.path.join(GOLDEN, "angles.npy")).astype(np.float32)
    return sino, angles


def _oracle_sino(sino, angles):
    from ccpi.reconstruction import instruments as ins
    setup = ins.diamond_setup(sino.shape, angles, 0.0, 1)
    g = setup.geometry
    b = np.zeros(g.sino_shape, np.float32)
    off = setup.data_v_offset
    b[:, :, off:off + sino.shape[1]] = -np.log(sino.transpose(0, 2, 1))
    og = oracle.Geometry(g.angles, g.h_pixels, g.v_pixels, g.origin, g.voxel_size, g.shape,
                         cone=False)
    return og, b


@pytest.mark.parametrize("variant,lam", [("cgls_tikhonov", 0.1), ("cgls_TVreg", 1e-4),
                                         ("cgls_TVreg", 5e-3)])
def test_parallel_regularised_on_reference_fixture(variant, lam):
    """The reference's fixture through parallelbeam.alg.cgls_tikhonov /
    cgls_TVreg (4 iterations), against ingest + oracle regularised CGLS +
    the final clamp; norm_r[0] != 0 is zeroed, norm_r[0] == 0 is left.
    At lambda 5e-3 the TV step is unstable on this data (float and FP64
    oracles part by 4 % of the max): bounded by that distance instead."""
    from ccpi.reconstruction.parallelbeam import alg
    sino, angles = _fixture()
    norm_r = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0], np.float32)
    vox = getattr(alg, variant)(sino, angles, np.double(0.0), 1, 4, 1, lam, norm_r, False)
    assert np.array_equal(norm_r, [0, 0, 0, 0, 5, 6])
    og, b = _oracle_sino(sino, angles)
    kind = oracle.TIKHONOV if variant == "cgls_tikhonov" else oracle.TV
    xo, _, no = oracle.cgls_reg(og, b, 4, kind, lam)
    xo = np.maximum(xo, 0)
    if lam < 1e-3 or kind == oracle.TIKHONOV:
        assert _rel(vox, xo) < 2e-4
    else:
        x64, _, _ = oracle.cgls_reg(og, b.astype(np.float64), 4, kind, lam, dtype=np.float64)
        x64 = np.maximum(x64, 0)
        assert _rel(vox, x64) <= 1.5 * _rel(xo, x64) + 1e-5
    assert np.allclose(alg.cgls_norms(), no, rtol=1e-3)
    keep = np.zeros(3, np.float32)
    keep[1] = 7.0
    getattr(alg, variant)(sino, angles, np.double(0.0), 1, 2, 1, lam, keep, False)
    assert np.array_equal(keep, [0, 7, 0])


def test_parallel_step_variants_continue_from_the_volume():
    """*_step adds the iterations' updates to last_iteration_volume, with the
    residual restarted from the data (cgls.cpp:45-49)."""
    from ccpi.reconstruction.parallelbeam import alg
    sino, angles = _fixture()
    og, b = _oracle_sino(sino, angles)
    prev = np.random.default_rng(1).random(og.shape).astype(np.float32) * 1e-3
    norm_r = np.zeros(2, np.float32)
    vox = alg.cgls_tikhonov_step(sino, angles, np.double(0.0), 1, 3, 1, 0.1, norm_r, False, prev)
    xo, _, _ = oracle.cgls_reg(og, b, 3, oracle.TIKHONOV, 0.1, x=prev)
    assert _rel(vox, np.maximum(xo, 0)) < 2e-4
    vox = alg.cgls_conv_step(sino, angles, np.double(0.0), 1, 3, 1, norm_r, False, prev)
    xo, _, _ = oracle.cgls(og, b, 3, x=prev)
    assert _rel(vox, np.maximum(xo, 0)) < 2e-4
    vox2 = alg.cgls_conv(sino, angles, np.double(0.0), 1, 3, 1, norm_r, False)
    xo2, _, _ = oracle.cgls(og, b, 3)
    assert _rel(vox2, np.maximum(xo2, 0)) < 2e-4
