"""Seeded synthetic workloads: geometries and ellipsoid phantoms.

Each workload fixes the geometry left open by its description (voxel size,
detector pitch, angle range), generates an additive-ellipsoid phantom from
``numpy.random.default_rng(seed)`` and measures it with the workload's
noise model.  Phantom voxels are filled on the device with torch.
"""

from __future__ import annotations

import math

import numpy as np

from .projectors import Geometry


def centred(n: int, pitch: float, shift: float = 0.0) -> np.ndarray:
    """Detector coordinates centred on 0 (Core/src/diamond.cpp:72-79)."""
    return (np.arange(n) - (n - 1) / 2.0) * pitch - shift


def cone_geometry(n: int, vox: float, n_angles: int, n_det: int, pitch: float,
                  sod: float, sdd: float) -> Geometry:
    """Cubic n^3 volume centred on the rotation axis, angles over [0, 2 pi),
    square centred flat detector, source on the x axis at z = 0."""
    ang = np.arange(n_angles) * (2.0 * math.pi / n_angles)
    o = -0.5 * n * vox
    return Geometry(ang, centred(n_det, pitch), centred(n_det, pitch), (o, o, o),
                    (vox, vox, vox), (n, n, n), cone=True, source=(-sod, 0.0, 0.0),
                    detector_x=sdd - sod)


def parallel_geometry(n: int, slices: int, n_angles: int, width: int,
                      pitch: float = 1.0, arc: float = math.pi) -> Geometry:
    """n x n x slices, 1 mm voxels, angles over [0, arc), centred detector of
    `width` columns and `slices` rows (ppv = 1)."""
    ang = np.arange(n_angles) * (arc / n_angles)
    return Geometry(ang, centred(width, pitch), centred(slices, 1.0),
                    (-0.5 * n, -0.5 * n, -0.5 * slices), (1.0, 1.0, 1.0), (n, n, slices))


# configuration table (sizes in voxels / pixels, lengths in mm)
CONE_512 = dict(n=512, vox=0.5, n_angles=720, n_det=768, pitch=1.0, sod=500.0, sdd=1000.0)
CONE_256 = dict(n=256, vox=1.0, n_angles=360, n_det=384, pitch=2.0, sod=500.0, sdd=1000.0)
CONE_768 = dict(n=768, vox=1.0 / 3.0, n_angles=1200, n_det=1024, pitch=0.75, sod=500.0,
                sdd=1000.0)
# the 768^3 precision study's scan at a third of its sampling (same 256 mm
# volume, 768 mm detector and distances): the size its FP64 oracle arm runs at
CONE_768_AT_256 = dict(n=256, vox=1.0, n_angles=400, n_det=342, pitch=2.25, sod=500.0,
                       sdd=1000.0)


def ellipsoid_params(n: int, seed: int, dens=(0.1, 1.0)):
    """Ellipsoid parameters in units of the volume half-extent: centre in the
    xy disc of radius 0.5 with |z| <= 0.5, semi-axes U[0.05, 0.35], a random
    rotation about z, density U[dens]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        r = 0.5 * math.sqrt(rng.random())
        t = 2 * math.pi * rng.random()
        c = (r * math.cos(t), r * math.sin(t), (rng.random() * 2 - 1) * 0.5)
        ax = tuple(rng.uniform(0.05, 0.35, 3))
        rot = 2 * math.pi * rng.random()
        d = rng.uniform(*dens)
        out.append((c, ax, rot, d))
    return out


def ellipsoid_phantom(geom: Geometry, n: int, seed: int, dens=(0.1, 1.0), device="cuda"):
    """Additive ellipsoids, voxel-centre inclusion, float32 on ``device``."""
    import torch
    nx, ny, nz = geom.shape
    vs, o = geom.voxel_size, geom.origin
    half = (0.5 * nx * vs[0], 0.5 * ny * vs[1], 0.5 * nz * vs[2])
    R = min(half[0], half[1])
    f64 = dict(dtype=torch.float64, device=device)
    cx = o[0] + (torch.arange(nx, **f64) + 0.5) * vs[0]
    cy = o[1] + (torch.arange(ny, **f64) + 0.5) * vs[1]
    cz = o[2] + (torch.arange(nz, **f64) + 0.5) * vs[2]
    vol = torch.zeros((nx, ny, nz), dtype=torch.float32, device=device)
    for c, ax, rot, d in ellipsoid_params(n, seed, dens):
        c0, c1, c2 = c[0] * R, c[1] * R, c[2] * half[2]
        a0, a1, a2 = ax[0] * R, ax[1] * R, ax[2] * half[2]
        cr, sr = math.cos(rot), math.sin(rot)
        X = (cx - c0)[:, None]
        Y = (cy - c1)[None, :]
        u = (X * cr + Y * sr) / a0
        w = (-X * sr + Y * cr) / a1
        q2 = u * u + w * w
        zz = ((cz - c2) / a2) ** 2
        vol += (q2[:, :, None] + zz[None, None, :] <= 1.0).to(torch.float32) * float(d)
    return vol


def ellipsoid_phantom_host(geom: Geometry, n: int, seed: int, dens=(0.1, 1.0)) -> np.ndarray:
    """The phantom of ``ellipsoid_phantom`` built with numpy on the host
    (float32 [nx][ny][nz]): each ellipsoid is evaluated on its bounding box
    only, with the same voxel-centre test and float32 additions in the same
    order.  Host-built inputs are identical on every machine, so results
    computed elsewhere (committed fixtures) refer to the same phantom."""
    nx, ny, nz = geom.shape
    vs, o = geom.voxel_size, geom.origin
    half = (0.5 * nx * vs[0], 0.5 * ny * vs[1], 0.5 * nz * vs[2])
    R = min(half[0], half[1])
    cx = o[0] + (np.arange(nx) + 0.5) * vs[0]
    cy = o[1] + (np.arange(ny) + 0.5) * vs[1]
    cz = o[2] + (np.arange(nz) + 0.5) * vs[2]
    vol = np.zeros((nx, ny, nz), np.float32)

    def span(centres, lo, hi):
        i0 = int(np.searchsorted(centres, lo, side="left"))
        i1 = int(np.searchsorted(centres, hi, side="right"))
        return max(i0 - 1, 0), min(i1 + 1, centres.size)

    for c, ax, rot, d in ellipsoid_params(n, seed, dens):
        c0, c1, c2 = c[0] * R, c[1] * R, c[2] * half[2]
        a0, a1, a2 = ax[0] * R, ax[1] * R, ax[2] * half[2]
        rxy = max(a0, a1)
        i0, i1 = span(cx, c0 - rxy, c0 + rxy)
        j0, j1 = span(cy, c1 - rxy, c1 + rxy)
        k0, k1 = span(cz, c2 - a2, c2 + a2)
        cr, sr = math.cos(rot), math.sin(rot)
        X = (cx[i0:i1] - c0)[:, None]
        Y = (cy[j0:j1] - c1)[None, :]
        u = (X * cr + Y * sr) / a0
        w = (-X * sr + Y * cr) / a1
        q2 = u * u + w * w
        zz = ((cz[k0:k1] - c2) / a2) ** 2
        m = q2[:, :, None] + zz[None, None, :] <= 1.0
        box = vol[i0:i1, j0:j1, k0:k1]
        box[m] += np.float32(d)
    return vol


def poisson_log_host(sino: np.ndarray, i0: float, seed: int) -> np.ndarray:
    """Host form of ``poisson_log`` with numpy's PCG64 stream: counts ~
    Poisson(I0 exp(-b)) drawn in C order (one angle at a time, which draws
    the same sequence as one call), then -ln(max(counts, 1) / I0) as
    float32."""
    rng = np.random.default_rng(seed)
    out = np.empty(sino.shape, np.float32)
    for a in range(sino.shape[0]):
        lam = i0 * np.exp(-sino[a].astype(np.float64))
        counts = rng.poisson(lam).astype(np.float64)
        out[a] = -np.log(np.maximum(counts, 1.0) / i0)
    return out


def poisson_log(sino, i0: float, seed: int):
    """Counts ~ Poisson(I0 exp(-b)); returns -ln(max(counts, 1) / I0)."""
    import torch
    g = torch.Generator(device=sino.device)
    g.manual_seed(seed)
    counts = torch.poisson(i0 * torch.exp(-sino.double()), generator=g)
    return (-torch.log(torch.clamp(counts, min=1.0) / i0)).float().contiguous()


def gaussian_noise(sino, rel_sigma: float, seed: int):
    import torch
    g = torch.Generator(device=sino.device)
    g.manual_seed(seed)
    sigma = rel_sigma * float(sino.max())
    return (sino + sigma * torch.randn(sino.shape, generator=g, device=sino.device)).contiguous()


def batched_problems(count: int = 64, seed: int = 3, sizes=(128, 256, 384, 512),
                     slices=(1, 32), angles=(90, 180, 360)):
    """Uneven parallel-beam problems: N in `sizes`, S in [slices[0],
    slices[1]], angle counts from `angles` over [0, pi), detector width
    ceil(sqrt(2) N), 1 mm pixels (one geometry per problem)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice(list(sizes)))
        s = int(rng.integers(slices[0], slices[1] + 1))
        na = int(rng.choice(list(angles)))
        out.append(parallel_geometry(n, s, na, int(math.ceil(math.sqrt(2.0) * n))))
    return out
