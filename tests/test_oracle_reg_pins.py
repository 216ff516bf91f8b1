"""Pins of the oracle's regularisation terms and regularised CGLS.

Tikhonov (Core/src/tikhonov.cpp:4-32) and TV (Core/src/total_v.cpp:21-195)
are pinned by closed forms that a plausible slip breaks:
  * constants map to 0; quadratics i^2 / j^2 give 0.5 in the interior and
    -(n - 1.5) on the upper face (mirror boundaries); a z-only field gives 0
    (Tikhonov does not couple z);
  * TV of a linear field a = al i + be j + ga k is 0 inside, -2 al / |g| on
    the upper x face, 0 on the lower faces, and on an upper x/y edge
    -2 al / sqrt(al^2 + ga^2) - 2 be / sqrt(be^2 + ga^2) (each denominator
    carries the min-mod terms of the other two axes); a slope below
    sqrt(FLT_EPSILON) shows the epsilon; TV(-a) = -TV(a); and for a radial
    cone sqrt(x^2 + y^2) the term approaches the curvature 1/r;
  * regularised CGLS (cgls.cpp:343-435): lambda = 0 is plain CGLS (itself
    pinned against lstsq), and the first two iterations follow from dense
    operators: the projector from exact intersections, Tikhonov from
    numpy's reflect padding.
"""

import math

import numpy as np
import pytest

import oracle
from geometries import ALL_CB, ALL_PB, dense, oracle_geom

FLT_EPS = float(np.finfo(np.float32).eps)


def _grid(shape):
    return np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_constant_fields_map_to_zero(dtype):
    a = np.full((5, 4, 3), 2.5, dtype)
    for kind in (oracle.TIKHONOV, oracle.TV):
        assert np.all(oracle.regularize(kind, a, dtype) == 0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tikhonov_quadratics_and_mirror_faces(dtype):
    nx, ny, nz = 6, 7, 3
    i, j, k = _grid((nx, ny, nz))
    t = oracle.regularize(oracle.TIKHONOV, (i * i).astype(dtype), dtype)
    assert np.all(t[:-1] == 0.5) and np.all(t[-1] == -(nx - 1.5))
    t = oracle.regularize(oracle.TIKHONOV, (j * j).astype(dtype), dtype)
    assert np.all(t[:, :-1] == 0.5) and np.all(t[:, -1] == -(ny - 1.5))
    t = oracle.regularize(oracle.TIKHONOV, (k * k + 1).astype(dtype), dtype)
    assert np.all(t == 0)


def _tikhonov_np(a):
    p = np.pad(a, ((1, 1), (1, 1), (0, 0)), mode="reflect")
    return 0.25 * (p[2:, 1:-1] + p[:-2, 1:-1] + p[1:-1, 2:] + p[1:-1, :-2]) - a


def test_tikhonov_matches_reflect_padding():
    a = np.random.default_rng(2).random((7, 5, 4))
    assert np.allclose(oracle.regularize(oracle.TIKHONOV, a, np.float64), _tikhonov_np(a),
                       rtol=0, atol=1e-15)
    a32 = a.astype(np.float32)
    assert np.allclose(oracle.regularize(oracle.TIKHONOV, a32), _tikhonov_np(a32.astype(np.float64)),
                       rtol=0, atol=2e-7)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tv_linear_field_closed_forms(dtype):
    al, be, ga = 1.0, 2.0, 3.0
    nx, ny, nz = 5, 6, 7
    i, j, k = _grid((nx, ny, nz))
    B = oracle.regularize(oracle.TV, (al * i + be * j + ga * k).astype(dtype), dtype).astype(np.float64)
    tol = 1e-6 if dtype == np.float32 else 1e-7
    g = math.sqrt(al * al + be * be + ga * ga)
    inner = (slice(1, -1),) * 3
    assert np.abs(B[inner]).max() < tol
    assert np.allclose(B[-1, 1:-1, 1:-1], -2 * al / g, atol=tol)
    assert np.allclose(B[1:-1, -1, 1:-1], -2 * be / g, atol=tol)
    assert np.allclose(B[1:-1, 1:-1, -1], -2 * ga / g, atol=tol)
    for face in (B[0, 1:-1, 1:-1], B[1:-1, 0, 1:-1], B[1:-1, 1:-1, 0]):
        assert np.abs(face).max() < tol
    edge = -2 * al / math.hypot(al, ga) - 2 * be / math.hypot(be, ga)
    assert np.allclose(B[-1, -1, 1:-1], edge, atol=tol)


def test_tv_epsilon_in_the_denominator():
    s = 1e-4
    i, j, k = _grid((3, 3, 6))
    B = oracle.regularize(oracle.TV, s * k, np.float64)
    want = -2 * s / math.sqrt(s * s + FLT_EPS)
    assert abs(B[1, 1, -1] - want) < 1e-9 * abs(want)


def test_tv_is_odd():
    a = np.random.default_rng(3).random((6, 5, 4)).astype(np.float32)
    assert np.array_equal(oracle.regularize(oracle.TV, -a), -oracle.regularize(oracle.TV, a))


def test_tv_of_radial_cone_is_its_curvature():
    n = 64
    i, j, k = _grid((n, n, 4))
    r = np.hypot(i - n / 2 + 0.5, j - n / 2 + 0.5)
    B = oracle.regularize(oracle.TV, r, np.float64)
    for rr in (10.0, 15.0, 20.0, 25.0):
        m = (np.abs(r - rr) < 0.5) & (k == 1)
        assert abs(B[m].mean() * rr - 1) < 0.03


@pytest.mark.parametrize("kind", [oracle.TIKHONOV, oracle.TV])
def test_cgls_reg_with_zero_lambda_is_cgls(kind):
    g = ALL_PB["pb_small"]()
    og = oracle_geom(g)
    rng = np.random.default_rng(4)
    b = og.fp(rng.random(g.shape)) + 0.01 * rng.random(g.sino_shape)
    x0, r0, n0 = oracle.cgls(og, b, 6, dtype=np.float64)
    x1, r1, n1 = oracle.cgls_reg(og, b, 6, kind, 0.0, dtype=np.float64)
    # the oracle's OpenMP reductions add in a varying order: two runs of the
    # same CGLS differ by ~1e-12 of the residual after six iterations
    assert np.allclose(x1, x0, rtol=0, atol=1e-10 * np.abs(x0).max())
    assert np.allclose(r1, r0, rtol=0, atol=1e-10 * np.abs(r0).max())
    assert np.allclose(n1, n0, rtol=1e-10)


def _tikhonov_matrix(shape):
    n = int(np.prod(shape))
    L = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1.0
        L[:, c] = _tikhonov_np(e.reshape(shape)).ravel()
    return L


@pytest.mark.parametrize("case", ["pb_small", "cb_small"])
def test_cgls_tikhonov_first_two_iterations(case):
    g = (ALL_PB if case in ALL_PB else ALL_CB)[case]()
    og = oracle_geom(g)
    A = dense(g)
    L = _tikhonov_matrix(g.shape)
    lam = 0.3
    rng = np.random.default_rng(8)
    b = A @ rng.random(A.shape[1])
    x, r, norms = oracle.cgls_reg(og, b.reshape(g.sino_shape), 2, oracle.TIKHONOV, lam,
                                  dtype=np.float64)
    # iteration 1 (cgls.cpp:362-401)
    d = A.T @ b
    gam = d @ d
    Ad = A @ d
    adv = lam * (L @ d)
    alpha = gam / (Ad @ Ad + adv @ adv)
    xe = alpha * d
    re = b - alpha * Ad
    xe = xe + lam * (L @ xe)
    s = A.T @ re
    n1 = math.sqrt(s @ s)
    d = s + (s @ s) / gam * d
    gam = s @ s
    # iteration 2: the voxel term keeps accumulating
    Ad = A @ d
    adv = adv + lam * (L @ d)
    alpha = gam / (Ad @ Ad + adv @ adv)
    xe = xe + alpha * d
    re = re - alpha * Ad
    xe = xe + lam * (L @ xe)
    s = A.T @ re
    scale = np.abs(xe).max()
    assert np.abs(x.ravel() - xe).max() < 1e-11 * scale
    assert np.abs(r.ravel() - re).max() < 1e-11 * np.abs(b).max()
    assert abs(norms[0] - n1) < 1e-11 * n1
    assert abs(norms[1] - math.sqrt(s @ s)) < 1e-11 * n1


def test_cgls_tv_first_iteration():
    g = ALL_PB["pb_small"]()
    og = oracle_geom(g)
    A = dense(g)
    lam = 0.05
    rng = np.random.default_rng(9)
    b = A @ rng.random(A.shape[1])
    x, _, norms = oracle.cgls_reg(og, b.reshape(g.sino_shape), 1, oracle.TV, lam,
                                  dtype=np.float64)
    tv = lambda v: oracle.regularize(oracle.TV, v.reshape(g.shape), np.float64).ravel()
    d = A.T @ b
    Ad = A @ d
    adv = lam * tv(d)
    alpha = (d @ d) / (Ad @ Ad + adv @ adv)
    xe = alpha * d
    xe = xe + lam * tv(xe)
    s = A.T @ (b - alpha * Ad)
    assert np.abs(x.ravel() - xe).max() < 1e-11 * np.abs(xe).max()
    assert abs(norms[0] - math.sqrt(s @ s)) < 1e-11 * norms[0]
