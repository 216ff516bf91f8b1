"""Independent float64 references used to pin the oracle and the HIP path.

Nothing here shares code with the oracle: every system-matrix entry is the
exact intersection length of a ray with one voxel, computed by the slab
method one voxel at a time (the method of the reference's single-voxel
check, Core/src/point.cpp:62-173), vectorised over voxels with numpy.
"""

from __future__ import annotations

import numpy as np


def _slab(p, d, lo, hi):
    """Parameter interval of p + t d inside [lo, hi) along one axis.

    Arrays broadcast; a direction component of 0 gives (-inf, inf) when p is
    inside the slab and an empty interval otherwise."""
    p, d, lo, hi = np.broadcast_arrays(*(np.asarray(q, np.float64) for q in (p, d, lo, hi)))
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (lo - p) / d
        t1 = (hi - p) / d
    tmin = np.minimum(t0, t1)
    tmax = np.maximum(t0, t1)
    flat = d == 0.0
    inside = (p >= lo) & (p < hi)
    tmin = np.where(flat, np.where(inside, -np.inf, np.inf), tmin)
    tmax = np.where(flat, np.where(inside, np.inf, -np.inf), tmax)
    return tmin, tmax


def parallel_matrix(angles, h_pix, v_pix, origin, vsize, shape):
    """Dense parallel-beam system matrix [(a,h,v), (i,j,k)] in float64.

    Ray (a, h) is the infinite line through (cos a * D - sin a * h,
    sin a * D + cos a * h) with direction (cos a, sin a); row v reads slice
    floor((v_pix[v] - z0) / dz).  Entries are exact chord lengths in mm."""
    nx, ny, nz = shape
    na, nh, nv = len(angles), len(h_pix), len(v_pix)
    xs = origin[0] + np.arange(nx + 1) * vsize[0]
    ys = origin[1] + np.arange(ny + 1) * vsize[1]
    X0, Y0 = np.meshgrid(xs[:-1], ys[:-1], indexing="ij")
    X1, Y1 = np.meshgrid(xs[1:], ys[1:], indexing="ij")
    kmap = np.floor((np.asarray(v_pix) - origin[2]) / vsize[2]).astype(int)
    A = np.zeros((na, nh, nv, nx, ny, nz))
    for a, phi in enumerate(angles):
        c, s = np.cos(phi), np.sin(phi)
        if abs(c) < 1e-12:
            c = 0.0
        if abs(s) < 1e-12:
            s = 0.0
        for h, hp in enumerate(h_pix):
            px, py = -s * hp, c * hp
            tx0, tx1 = _slab(px, c, X0, X1)
            ty0, ty1 = _slab(py, s, Y0, Y1)
            L = np.maximum(0.0, np.minimum(tx1, ty1) - np.maximum(tx0, ty0))
            L[~np.isfinite(L)] = 0.0
            for v in range(nv):
                if 0 <= kmap[v] < nz:
                    A[a, h, v, :, :, kmap[v]] = L
    return A.reshape(na * nh * nv, nx * ny * nz)


def cone_matrix(source, det_x, angles, h_pix, v_pix, origin, vsize, shape):
    """Dense cone-beam system matrix in float64: exact 3-D length of the
    source -> pixel segment inside each voxel (mm).  Source (sx, sy, sz) and
    pixel (det_x, h, v) are rotated about z by each angle."""
    nx, ny, nz = shape
    sx, sy, sz = source
    na, nh, nv = len(angles), len(h_pix), len(v_pix)
    g = [origin[q] + np.arange(n + 1) * vsize[q] for q, n in enumerate(shape)]
    X0, Y0, Z0 = np.meshgrid(g[0][:-1], g[1][:-1], g[2][:-1], indexing="ij")
    X1, Y1, Z1 = np.meshgrid(g[0][1:], g[1][1:], g[2][1:], indexing="ij")
    A = np.zeros((na, nh, nv, nx, ny, nz))
    for a, phi in enumerate(angles):
        c, s = np.cos(phi), np.sin(phi)
        p1 = np.array([c * sx - s * sy, s * sx + c * sy, sz])
        for h, hp in enumerate(h_pix):
            for v, vp in enumerate(v_pix):
                p2 = np.array([c * det_x - s * hp, s * det_x + c * hp, vp])
                d = p2 - p1
                d[np.abs(d) < 1e-12] = 0.0
                tx0, tx1 = _slab(p1[0], d[0], X0, X1)
                ty0, ty1 = _slab(p1[1], d[1], Y0, Y1)
                tz0, tz1 = _slab(p1[2], d[2], Z0, Z1)
                t0 = np.maximum(np.maximum(np.maximum(tx0, ty0), tz0), 0.0)
                t1 = np.minimum(np.minimum(np.minimum(tx1, ty1), tz1), 1.0)
                L = np.maximum(0.0, t1 - t0) * np.linalg.norm(p2 - p1)
                A[a, h, v] = L
    return A.reshape(na * nh * nv, nx * ny * nz)


def cgls_textbook(A, b, niter):
    """CGLS (Bjorck, Numerical Methods for Least Squares, alg. 7.4.1) from
    x0 = 0 in float64; returns iterates and ||A^T r_k||."""
    x = np.zeros(A.shape[1])
    r = b.astype(np.float64).copy()
    s = A.T @ r
    d = s.copy()
    gamma = s @ s
    xs, norms = [], []
    for _ in range(niter):
        q = A @ d
        alpha = gamma / (q @ q)
        x = x + alpha * d
        r = r - alpha * q
        s = A.T @ r
        gnew = s @ s
        d = s + (gnew / gamma) * d
        gamma = gnew
        xs.append(x.copy())
        norms.append(np.sqrt(gamma))
    return xs, np.array(norms)


def cone_box_lengths(source, det_x, angle, h_pix, v_pix, boxes):
    """Exact length (mm) inside a set of disjoint axis-aligned boxes
    ((lo_x, lo_y, lo_z), (hi_x, hi_y, hi_z)) of every source -> pixel
    segment of one angle, as [nh][nv] in float64 (same ray convention as
    cone_matrix)."""
    c, s = np.cos(angle), np.sin(angle)
    sx, sy, sz = source
    p1 = (c * sx - s * sy, s * sx + c * sy, sz)
    H, V = np.meshgrid(np.asarray(h_pix, np.float64), np.asarray(v_pix, np.float64),
                       indexing="ij")
    p2 = (c * det_x - s * H, s * det_x + c * H, V)
    d = [p2[q] - p1[q] for q in range(3)]
    for q in range(3):
        d[q] = np.where(np.abs(d[q]) < 1e-12, 0.0, d[q])
    norm = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
    total = np.zeros(H.shape)
    for lo, hi in boxes:
        t = [_slab(p1[q], d[q], lo[q], hi[q]) for q in range(3)]
        t0 = np.maximum(np.maximum(np.maximum(t[0][0], t[1][0]), t[2][0]), 0.0)
        t1 = np.minimum(np.minimum(np.minimum(t[0][1], t[1][1]), t[2][1]), 1.0)
        total += np.maximum(0.0, t1 - t0) * norm
    return total
