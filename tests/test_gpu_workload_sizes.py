"""Batched parallel CGLS and the differentiable cone reconstruction at their
workload sizes (GPU test).

Batched parallel-beam CGLS, the 64 problems of the batched workload
(workloads.batched_problems(64, seed=3): N in {128, 256, 384, 512},
S in [1, 32], 90 / 180 / 360 angles, width ceil(sqrt2 N), 1 % Gaussian
noise), 20 iterations:
  * the batched form (one persistent launch per step, graph replay) against
    the same problems one at a time through ccpi_cgls: iterates within
    1e-6 of the maximum and norms within 1e-6 of the first, per problem
    (the batched items run the one-problem kernels' bodies, and its tiled
    backprojector items add the one-problem launch's angle groups in the
    same order, so the floats are the same; 20 iterations would amplify a
    different summation order to ~1e-4);
  * the adjoint identity <A x, y> = <x, A^T y> of every problem's projector;
  * the N = 128 problems against the FP64 oracle (the same model in double,
    cgls.cpp:26-157): the device iterate is no further from it than the
    float oracle's (the reference's arithmetic) is, within 1 % of that
    distance per problem and on average over the problems.  Both are float
    evaluations of 20 noisy-data iterations, 2.5 - 5 % from the FP64 iterate;
    another summation order of the same float terms moves that distance by
    a few tenths of a percent either way.

Differentiable cone reconstruction, 256^3 of 1 mm, 360 projections over
360 deg, 384^2 detector of 2 mm, SOD 500, SDD 1000: loss
0.5 ||A x - b||^2 + 1e-3 TV_iso(x) at a seeded random volume.
  * The data term is quadratic, so its gradient at voxel q is exactly
    <r, A e_q> with r = A x - b: A e_q is computed here in float64 as the
    exact length of every source -> pixel segment inside voxel q (slab
    method, the pixels of q's footprint per angle) and compared with the
    autograd gradient (hand-written backprojection) on 64 seeded voxels, to
    2e-5 of sum |r| A e_q.  The geometry's rays cross at most one z plane
    per xy cell (slope <= 0.384, 1.41 mm per cell diagonal against 1 mm
    slices), where the projector's weights are the exact lengths.
  * The TV term: its autograd gradient against float64 central differences
    (eps = 1e-3) on the same voxels, evaluated on the 5^3 block around q
    (the only terms that depend on x_q).
"""

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch64():
    from ccpi.reconstruction import workloads as W
    from ccpi.reconstruction.projectors import forward_project
    geoms = W.batched_problems(64, seed=3)
    bs = []
    for i, g in enumerate(geoms):
        ph = W.ellipsoid_phantom(g, 6, seed=300 + i, dens=(0.1, 1.0))
        bs.append(W.gaussian_noise(forward_project(g, ph), 0.01, seed=300 + i))
    return geoms, bs


def test_batched_64_matches_one_at_a_time(batch64):
    from ccpi.reconstruction.projectors import cgls, cgls_batched
    geoms, bs = batch64
    print("\nbatched 64 problems, 20 iterations", flush=True)
    xb, _, nb = cgls_batched(geoms, bs, 20)
    worst = 0.0
    for i, (g, b) in enumerate(zip(geoms, bs)):
        xp, _, npl = cgls(g, b, 20)
        dev = float((xb[i] - xp).abs().max() / xp.abs().max())
        worst = max(worst, dev)
        # ||A^T r|| falls by orders of magnitude over the run, towards the
        # level of its float rounding: compared on the scale of its value
        # after the first iteration
        nb_i, np_i = np.asarray(nb[i]), np.asarray(npl)
        ndev = float(np.max(np.abs(nb_i - np_i)) / np_i[0])
        print(f"problem {i} {g.shape}: iterate {dev:.2e}, norms {ndev:.2e} of ||A^T b||, "
              f"last norm {np_i[-1] / np_i[0]:.2e} of it", flush=True)
        assert dev <= 1e-6, (i, g.shape, dev)
        assert ndev <= 1e-6, (i, g.shape, ndev)
    print(f"max deviation batched vs one at a time {worst:.2e}", flush=True)


def test_batched_64_adjoint(batch64):
    from ccpi.reconstruction.projectors import backward_project, forward_project
    geoms, _ = batch64
    gen = torch.Generator(device="cuda").manual_seed(31)
    for i, g in enumerate(geoms):
        x = torch.rand(g.shape, device="cuda", generator=gen)
        y = torch.rand(g.sino_shape, device="cuda", generator=gen)
        lhs = float(torch.sum(forward_project(g, x).double() * y.double()))
        rhs = float(torch.sum(x.double() * backward_project(g, y).double()))
        assert abs(lhs - rhs) <= 1e-5 * abs(lhs), (i, g.shape, lhs, rhs)


def test_batched_128_problems_against_fp64_oracle(batch64):
    from ccpi.reconstruction.projectors import cgls_batched
    geoms, bs = batch64
    idx = [i for i, g in enumerate(geoms) if g.shape[0] == 128]
    assert len(idx) >= 4
    xb, _, _ = cgls_batched([geoms[i] for i in idx], [bs[i] for i in idx], 20)
    devs, refs = [], []
    for q, i in enumerate(idx):
        g = geoms[i]
        og = oracle.Geometry(g.angles, g.h_pixels, g.v_pixels, g.origin, g.voxel_size,
                             g.shape)
        b = bs[i].cpu().numpy()
        x64, _, _ = oracle.cgls(og, b.astype(np.float64), 20, dtype=np.float64)
        x32, _, _ = oracle.cgls(og, b, 20)
        nrm = np.linalg.norm(x64)
        e_dev = np.linalg.norm(xb[q].cpu().numpy() - x64) / nrm
        e_ref = np.linalg.norm(x32 - x64) / nrm
        print(f"problem {i} {g.shape} {g.angles.size} angles: distance to FP64 "
              f"device {e_dev:.3e} float oracle {e_ref:.3e}", flush=True)
        assert e_dev <= 1.01 * e_ref
        devs.append(e_dev)
        refs.append(e_ref)
    print(f"mean distance to FP64: device {np.mean(devs):.4e} float oracle "
          f"{np.mean(refs):.4e}", flush=True)
    assert np.mean(devs) <= np.mean(refs)


# ---- differentiable cone reconstruction at 256^3 ---------------------------

def _voxel_column(g, r, q):
    """<r, A e_q> and sum |r| A e_q in float64 for voxel index q = (i, j, k)
    from the exact segment lengths inside the voxel."""
    i, j, k = q
    vs, o = g.voxel_size, g.origin
    lo = np.array([o[0] + i * vs[0], o[1] + j * vs[1], o[2] + k * vs[2]])
    hi = lo + np.array(vs)
    corners = np.array([[lo[0] if a == 0 else hi[0], lo[1] if b == 0 else hi[1],
                         lo[2] if c == 0 else hi[2]]
                        for a in (0, 1) for b in (0, 1) for c in (0, 1)])
    sx, sy, sz = g.source
    hp, vp = g.h_pixels, g.v_pixels
    dot = 0.0
    mag = 0.0
    for a, phi in enumerate(g.angles):
        c, s = math.cos(phi), math.sin(phi)
        p1 = np.array([c * sx - s * sy, s * sx + c * sy, sz])
        # detector frame: u along the source -> detector axis, w along h
        ux, uy = c, s
        wx, wy = -s, c
        rel = corners - p1
        depth = rel[:, 0] * ux + rel[:, 1] * uy
        L = g.detector_x - sx
        hh = (rel[:, 0] * wx + rel[:, 1] * wy) * L / depth + sy
        vv = rel[:, 2] * L / depth + sz
        h0 = max(int(np.searchsorted(hp, hh.min())) - 2, 0)
        h1 = min(int(np.searchsorted(hp, hh.max())) + 2, hp.size)
        v0 = max(int(np.searchsorted(vp, vv.min())) - 2, 0)
        v1 = min(int(np.searchsorted(vp, vv.max())) + 2, vp.size)
        if h0 >= h1 or v0 >= v1:
            continue
        H, V = np.meshgrid(hp[h0:h1], vp[v0:v1], indexing="ij")
        p2 = (c * g.detector_x - s * H, s * g.detector_x + c * H, V)
        d = [p2[t] - p1[t] for t in range(3)]
        t0 = np.zeros(H.shape)
        t1 = np.ones(H.shape)
        for t in range(3):
            with np.errstate(divide="ignore", invalid="ignore"):
                ta = (lo[t] - p1[t]) / d[t]
                tb = (hi[t] - p1[t]) / d[t]
            t0 = np.maximum(t0, np.minimum(ta, tb))
            t1 = np.minimum(t1, np.maximum(ta, tb))
        length = np.maximum(0.0, t1 - t0) * np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
        rr = r[a, h0:h1, v0:v1]
        dot += float(np.sum(rr * length))
        mag += float(np.sum(np.abs(rr) * length))
    return dot, mag


@pytest.fixture(scope="module")
def cone256():
    from ccpi.reconstruction import workloads as W
    from ccpi.reconstruction.projectors import forward_project
    g = W.cone_geometry(**W.CONE_256)
    ph = torch.from_numpy(W.ellipsoid_phantom_host(g, 30, seed=4, dens=(0.0, 0.05))).cuda()
    b = W.poisson_log(forward_project(g, ph), 1e4, seed=4)
    gen = torch.Generator(device="cuda").manual_seed(41)
    x = torch.rand(g.shape, device="cuda", generator=gen, dtype=torch.float32)
    rng = np.random.default_rng(42)
    vox = [tuple(int(t) for t in rng.integers(2, 254, 3)) for _ in range(64)]
    return g, b, x, vox


def test_cone256_data_gradient_is_the_exact_column_product(cone256):
    from ccpi.reconstruction.autograd import data_fidelity, project
    g, b, x, vox = cone256
    print("\n256^3 data-term gradient on 64 voxels", flush=True)
    xt = x.clone().requires_grad_(True)
    loss = data_fidelity(xt, g, b)
    grad, = torch.autograd.grad(loss, xt)
    with torch.no_grad():
        r = (project(x, g) - b).double().cpu().numpy()
    gq = grad.double().cpu().numpy()
    worst = 0.0
    for q in vox:
        dot, mag = _voxel_column(g, r, q)
        err = abs(gq[q] - dot) / mag
        worst = max(worst, err)
        assert err <= 2e-5, (q, gq[q], dot, mag)
    print(f"max |grad - <r, A e_q>| / sum |r| A e_q = {worst:.2e}", flush=True)


def test_cone256_tv_gradient_matches_central_differences(cone256):
    from ccpi.reconstruction.autograd import tv_isotropic
    g, _, x, vox = cone256
    x64 = x.double().clone().requires_grad_(True)
    lam = 1e-3
    tv = lam * tv_isotropic(x64)
    grad, = torch.autograd.grad(tv, x64)
    xc = x.double()
    eps = 1e-3
    for (i, j, k) in vox:
        blk = xc[i - 2:i + 3, j - 2:j + 3, k - 2:k + 3].clone()
        bp, bm = blk.clone(), blk.clone()
        bp[2, 2, 2] += eps
        bm[2, 2, 2] -= eps
        fd = float(lam * (tv_isotropic(bp) - tv_isotropic(bm)) / (2 * eps))
        an = float(grad[i, j, k])
        assert abs(an - fd) <= 1e-4 * max(abs(fd), lam), ((i, j, k), an, fd)
