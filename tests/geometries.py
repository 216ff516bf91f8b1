"""Small seeded test geometries shared by the CPU and GPU test suites."""

from __future__ import annotations

import numpy as np


def centred(n, pitch, shift=0.0):
    """Detector coordinates centred on 0 (Core/src/diamond.cpp:72-79)."""
    return (np.arange(n) - (n - 1) / 2.0) * pitch - shift


class PB:
    """Parallel-beam case: [x][y][z] volume, [angle][h][v] sinogram."""

    def __init__(self, shape, vsize, angles, h_pix, v_pix, origin=None):
        self.cone = False
        self.shape = tuple(shape)
        self.vsize = np.asarray(vsize, np.float64)
        self.origin = (-self.vsize * np.asarray(shape) / 2.0 if origin is None
                       else np.asarray(origin, np.float64))
        self.angles = np.asarray(angles, np.float64)
        self.h_pix = np.asarray(h_pix, np.float64)
        self.v_pix = np.asarray(v_pix, np.float64)
        self.source = (0.0, 0.0, 0.0)
        self.det_x = 0.0

    @property
    def sino_shape(self):
        return (self.angles.size, self.h_pix.size, self.v_pix.size)


class CB(PB):
    def __init__(self, shape, vsize, angles, h_pix, v_pix, sod, sdd, origin=None,
                 source_y=0.0, source_z=0.0):
        super().__init__(shape, vsize, angles, h_pix, v_pix, origin)
        self.cone = True
        self.source = (-float(sod), float(source_y), float(source_z))
        self.det_x = float(sdd) - float(sod)


def pb_small():
    """16x16x4, 12 angles incl. 0/45/90/135 deg, 24 columns, COR shift."""
    ang = np.deg2rad([0, 13, 45, 61, 90, 97.5, 135, 150, 180, 211, 270, 333.3])
    return PB((16, 16, 4), (1.0, 1.0, 1.0), ang, centred(24, 1.0, 0.37),
              centred(4, 1.0))


def pb_ppv2():
    """ppv = 2: two detector rows per slice, non-square pixels."""
    ang = np.deg2rad(np.linspace(0, 180, 10, endpoint=False) + 3.0)
    return PB((10, 12, 3), (1.0, 1.0, 2.0), ang, centred(18, 0.9), centred(6, 1.0))


def pb_odd_tiles():
    """Grid larger than one 32x32 tile in x and y, ragged last tile."""
    ang = np.deg2rad([0.0, 22.5, 45.0, 89.0, 90.0, 123.0, 179.0])
    return PB((40, 35, 2), (0.5, 0.5, 1.0), ang, centred(60, 0.5, 0.11), centred(2, 1.0))


def cb_small():
    """12^3, 16 angles over 360 deg, 20x20 detector, SOD 40, SDD 80."""
    ang = np.linspace(0, 2 * np.pi, 16, endpoint=False) + 0.05
    return CB((12, 12, 12), (1.0, 1.0, 1.0), ang, centred(20, 1.6), centred(20, 1.6),
              40.0, 80.0)


def cb_clip():
    """Detector taller than the volume shadow: rays enter and leave through
    the top/bottom faces (the reference's ncom == 0 case)."""
    ang = np.linspace(0, 2 * np.pi, 10, endpoint=False) + 0.2
    return CB((10, 10, 6), (1.0, 1.0, 1.0), ang, centred(16, 1.5), centred(14, 1.5),
              30.0, 60.0)


def cb_offcentre():
    """Shifted source (XTek find_centre style), non-cubic voxels."""
    ang = np.linspace(0, 2 * np.pi, 12, endpoint=False)
    return CB((9, 11, 8), (0.8, 0.8, 1.0), ang, centred(18, 1.2, 0.4), centred(12, 1.3),
              35.0, 70.0, source_y=0.35)


ALL_PB = {"pb_small": pb_small, "pb_ppv2": pb_ppv2, "pb_odd_tiles": pb_odd_tiles}
ALL_CB = {"cb_small": cb_small, "cb_clip": cb_clip, "cb_offcentre": cb_offcentre}


def oracle_geom(g):
    import oracle
    return oracle.Geometry(g.angles, g.h_pix, g.v_pix, g.origin, g.vsize, g.shape,
                           cone=g.cone, source=g.source, det_x=g.det_x)


def dense(g):
    import bruteforce as bf
    if g.cone:
        return bf.cone_matrix(g.source, g.det_x, g.angles, g.h_pix, g.v_pix, g.origin,
                              g.vsize, g.shape)
    return bf.parallel_matrix(g.angles, g.h_pix, g.v_pix, g.origin, g.vsize, g.shape)


def ellipsoid_phantom(shape, vsize, origin, n, seed, dens=(0.1, 1.0)):
    """Additive ellipsoids inside the inscribed cylinder, voxel-centre
    inclusion (the synthetic phantom family used by every workload)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    half = np.array([nx * vsize[0], ny * vsize[1], nz * vsize[2]]) / 2.0
    cx = origin[0] + (np.arange(nx) + 0.5) * vsize[0]
    cy = origin[1] + (np.arange(ny) + 0.5) * vsize[1]
    cz = origin[2] + (np.arange(nz) + 0.5) * vsize[2]
    X, Y = np.meshgrid(cx, cy, indexing="ij")
    vol = np.zeros(shape, np.float32)
    R = min(half[0], half[1])
    for _ in range(n):
        r = 0.5 * R * np.sqrt(rng.random())
        t = 2 * np.pi * rng.random()
        c = np.array([r * np.cos(t), r * np.sin(t), (rng.random() * 2 - 1) * 0.5 * half[2]])
        ax = rng.uniform(0.05, 0.35, 3) * np.array([R, R, half[2]])
        rot = 2 * np.pi * rng.random()
        d = rng.uniform(*dens)
        cr, sr = np.cos(rot), np.sin(rot)
        u = ((X - c[0]) * cr + (Y - c[1]) * sr) / ax[0]
        w = (-(X - c[0]) * sr + (Y - c[1]) * cr) / ax[1]
        q2 = u * u + w * w
        for k in range(nz):
            zz = (cz[k] - c[2]) / ax[2]
            m = q2 + zz * zz <= 1.0
            vol[:, :, k][m] += d
    return vol
