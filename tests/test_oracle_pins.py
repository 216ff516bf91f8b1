"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no value-level tests for this path and cannot be
compiled here (it needs Boost), so the oracle is pinned by checks that do
not restate its own formulas:
  * exact line/voxel intersection lengths computed voxel by voxel
    (tests/bruteforce.py) for FP and, transposed, for BP;
  * axis-aligned closed forms (p2D.cpp:233-283: weights are the voxel size);
  * the adjoint identity <Ax, y> = <x, A^T y>;
  * CGLS: least-squares limit (numpy.linalg.lstsq) and the one-step
    steepest-descent closed form; SIRT: convergence to the exact solution of
    a consistent system; MLEM: the sum(A x) = sum(b) invariant.
A deliberately wrong oracle (dropped d_conv, transposed h/v, wrong sign in
the z split) fails the brute-force comparisons by orders of magnitude more
than their tolerances.
"""

import numpy as np
import pytest

import oracle
from geometries import ALL_CB, ALL_PB, CB, PB, centred, dense, oracle_geom

# float32 weights come from float alpha differences; relative to the
# output scale the model differs from exact geometry by a few 1e-6.
F32_TOL = 2e-5
F64_TOL = 1e-11


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


@pytest.mark.parametrize("name", list(ALL_PB) + list(ALL_CB))
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fp_matches_exact_intersections(name, dtype):
    g = (ALL_PB | ALL_CB)[name]()
    og = oracle_geom(g)
    A = dense(g)
    rng = np.random.default_rng(7)
    x = rng.random(g.shape).astype(dtype)
    y = og.fp(x)
    yb = (A @ x.astype(np.float64).ravel()).reshape(y.shape)
    assert _rel(y, yb) < (F32_TOL if dtype == np.float32 else F64_TOL)


@pytest.mark.parametrize("name", list(ALL_PB) + list(ALL_CB))
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_bp_matches_exact_transpose(name, dtype):
    g = (ALL_PB | ALL_CB)[name]()
    og = oracle_geom(g)
    A = dense(g)
    rng = np.random.default_rng(8)
    y = rng.random(g.sino_shape).astype(dtype)
    x = og.bp(y)
    xb = (A.T @ y.astype(np.float64).ravel()).reshape(x.shape)
    assert _rel(x, xb) < (F32_TOL if dtype == np.float32 else F64_TOL)


def test_fp_accumulates_and_bp_accumulates():
    g = ALL_PB["pb_small"]()
    og = oracle_geom(g)
    rng = np.random.default_rng(1)
    x = rng.random(g.shape).astype(np.float64)
    y0 = rng.random(g.sino_shape)
    y = og.fp(x, y0.copy())
    assert np.allclose(y, y0 + og.fp(x), rtol=1e-12, atol=1e-12)
    x0 = rng.random(g.shape)
    xs = og.bp(y0, x0.copy())
    assert np.allclose(xs, x0 + og.bp(y0), rtol=1e-12, atol=1e-12)


def test_cone_fp_scales_prior_content_by_dconv():
    """c2D.cpp:616-623: the whole output slab, prior content included, is
    multiplied by the source->pixel distance."""
    g = ALL_CB["cb_small"]()
    og = oracle_geom(g)
    y0 = np.ones(g.sino_shape)
    y = og.fp(np.zeros(g.shape), y0.copy())
    hh, vv = np.meshgrid(g.h_pix, g.v_pix, indexing="ij")
    sx, sy, sz = g.source
    dconv = np.sqrt((g.det_x - sx) ** 2 + (hh - sy) ** 2 + (vv - sz) ** 2)
    assert np.allclose(y, np.broadcast_to(dconv, y.shape), rtol=1e-14)


@pytest.mark.parametrize("deg", [0.0, 90.0, 180.0, 270.0])
def test_parallel_axis_aligned_closed_form(deg):
    """At 0/90/180/270 degrees every crossed cell weighs one voxel width, so
    each ray is the voxel size times the sum along one grid line."""
    g = PB((8, 6, 2), (0.7, 0.7, 1.0), np.deg2rad([deg]), centred(14, 0.7, 0.05),
           centred(2, 1.0))
    og = oracle_geom(g)
    rng = np.random.default_rng(3)
    x = rng.random(g.shape)
    y = og.fp(x)[0]
    c, s = round(np.cos(np.deg2rad(deg))), round(np.sin(np.deg2rad(deg)))
    for h, hp in enumerate(g.h_pix):
        px, py = -s * hp, c * hp        # a point on the ray
        if c != 0:
            j = int(np.floor((py - g.origin[1]) / g.vsize[1]))
            ref = x[:, j, :].sum(0) * g.vsize[0] if 0 <= j < g.shape[1] else 0.0
        else:
            i = int(np.floor((px - g.origin[0]) / g.vsize[0]))
            ref = x[i, :, :].sum(0) * g.vsize[1] if 0 <= i < g.shape[0] else 0.0
        assert np.allclose(y[h], ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name", list(ALL_PB) + list(ALL_CB))
def test_adjoint_identity(name):
    g = (ALL_PB | ALL_CB)[name]()
    og = oracle_geom(g)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(g.shape)
    y = rng.standard_normal(g.sino_shape)
    lhs = np.vdot(og.fp(x), y)
    rhs = np.vdot(x, og.bp(y))
    assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + np.abs(og.fp(np.abs(x))).sum())


def _ls_case():
    ang = np.deg2rad(np.linspace(0, 180, 24, endpoint=False) + 1.0)
    return PB((6, 6, 1), (1.0, 1.0, 1.0), ang, centred(10, 0.8), centred(1, 1.0))


def test_cgls_reaches_least_squares_solution():
    g = _ls_case()
    og = oracle_geom(g)
    A = dense(g)
    rng = np.random.default_rng(5)
    b = A @ rng.random(A.shape[1]) + 0.01 * rng.standard_normal(A.shape[0])
    x, r, norms = oracle.cgls(og, b.reshape(g.sino_shape), 60, dtype=np.float64)
    xls = np.linalg.lstsq(A, b, rcond=None)[0]
    assert _rel(x.ravel(), xls) < 1e-8
    # the residual returned is b - A x
    assert np.allclose(r.ravel(), b - A @ x.ravel(), atol=1e-10)
    assert norms[-1] < 1e-8 * norms[0]


def test_cgls_first_step_is_steepest_descent():
    g = ALL_CB["cb_small"]()
    og = oracle_geom(g)
    A = dense(g)
    rng = np.random.default_rng(6)
    b = A @ rng.random(A.shape[1])
    x, _, norms = oracle.cgls(og, b.reshape(g.sino_shape), 1, dtype=np.float64)
    g0 = A.T @ b
    step = (g0 @ g0) / np.sum((A @ g0) ** 2)
    assert _rel(x.ravel(), step * g0) < 1e-10
    x32, _, _ = oracle.cgls(og, b.reshape(g.sino_shape).astype(np.float32), 1)
    assert _rel(x32.ravel(), step * g0) < 1e-4


def test_sirt_converges_to_consistent_solution():
    g = _ls_case()
    og = oracle_geom(g)
    A = dense(g)
    rng = np.random.default_rng(9)
    xt = 0.2 + rng.random(A.shape[1])
    b = A @ xt
    x = oracle.sirt(og, b.reshape(g.sino_shape), 4000, dtype=np.float64)
    assert _rel(x.ravel(), xt) < 1e-6


def test_mlem_preserves_total_counts_and_converges():
    g = _ls_case()
    og = oracle_geom(g)
    A = dense(g)
    rng = np.random.default_rng(10)
    xt = 0.2 + rng.random(A.shape[1])
    b = A @ xt
    for n in (1, 2, 5):
        x = oracle.mlem(og, b.reshape(g.sino_shape), n, dtype=np.float64)
        assert abs((A @ x.ravel()).sum() - b.sum()) < 1e-10 * b.sum()
    x = oracle.mlem(og, b.reshape(g.sino_shape), 3000, dtype=np.float64)
    assert _rel(x.ravel(), xt) < 1e-3


def _offz_case(sz):
    return CB((8, 8, 4), (1.0, 1.0, 1.0), np.linspace(0, 2 * np.pi, 10, endpoint=False),
              centred(16, 1.6), centred(40, 0.4), 40.0, 80.0, source_z=sz)


@pytest.mark.parametrize("sz", [0.3, 0.99, -0.99, -1.0, 1.0, 1.7, -1.7])
def test_source_height_fp_exact_bp_transpose_outside_end_slices(sz):
    """A source off the volume's mid-height (z range [-2, 2], 1 mm slices):
    A 1 is the exact ray-box chord (tests/bruteforce.py) for every height,
    and b2D is the transpose of f2D for heights in [-1, 1), the source
    outside the end slices.  From the top slice's lower plane up, or inside
    the bottom slice, the reference's b2D departs from the transpose
    (recorded here: 15 % at 1.0, a factor 12 at 1.3; the device BP
    reproduces it, tests/test_gpu_bp_paths.py)."""
    import bruteforce as bf
    g = _offz_case(sz)
    og = oracle_geom(g)
    y1 = og.fp(np.ones(g.shape))
    lo = tuple(g.origin)
    hi = tuple(g.origin[q] + g.shape[q] * g.vsize[q] for q in range(3))
    for a in range(len(g.angles)):
        want = bf.cone_box_lengths(g.source, g.det_x, g.angles[a], g.h_pix, g.v_pix, [(lo, hi)])
        assert np.abs(y1[a] - want).max() <= 1e-9 * want.max()
    rng = np.random.default_rng(11)
    x = rng.random(g.shape)
    y = rng.random(g.sino_shape)
    lhs = np.vdot(og.fp(x), y)
    rhs = np.vdot(x, og.bp(y))
    if -1.0 <= sz < 1.0:
        assert abs(lhs - rhs) <= 1e-11 * abs(lhs)
    else:
        assert abs(lhs - rhs) > 0.1 * abs(lhs)


@pytest.mark.parametrize("dy", [1.0, 1.2])
def test_parallel_bp_flat_top_rule_for_rectangular_voxels(dy):
    """The parallel b2D's trapezoid plateau is vox_size[0] / |cos| when
    |cos| > |sin|, else vox_size[1] / |sin| (p2D.cpp:1045-1048).  The
    longest chord of a dx x dy rectangle is min(dx / |cos|, dy / |sin|),
    which that rule picks only when dx == dy: with square voxels the FP64
    oracle's b2D is the exact transpose of the brute-force matrix, with
    dy = 1.2 dx it departs from it (and from the f2D, whose walk lengths
    are exact) by half a percent in the adjoint identity.  The device BP
    reproduces the rule (tests/test_gpu_par_bp_tiled.py compares it with
    the oracle; its adjoint checks use square voxels)."""
    import bruteforce as bf
    g = PB((6, 5, 2), (1.0, dy, 1.0), np.deg2rad([10.0, 35.0, 50.0, 80.0, 125.0, 160.0]),
           centred(14, 0.7, 0.1), centred(2, 1.0))
    og = oracle_geom(g)
    A = bf.parallel_matrix(g.angles, g.h_pix, g.v_pix, g.origin, g.vsize, g.shape)
    rng = np.random.default_rng(14)
    y = rng.random(g.sino_shape)
    x = rng.random(g.shape)
    bp = og.bp(y).ravel()
    want = A.T @ y.ravel()
    lhs = np.vdot(og.fp(x), y)
    rhs = np.vdot(x, og.bp(y))
    if dy == 1.0:
        assert np.abs(bp - want).max() <= 1e-9 * np.abs(want).max()
        assert abs(lhs - rhs) <= 1e-11 * abs(lhs)
    else:
        assert np.abs(bp - want).max() > 1e-3 * np.abs(want).max()
        assert abs(lhs - rhs) > 1e-3 * abs(lhs)
