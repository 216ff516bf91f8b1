"""Instrument set-up and scan ingest of the Python path.

Restates what the reference does between the Python call and the solver:
  Diamond (parallel beam)  setup_experimental_geometry(numpy)
                           Core/src/diamond.cpp:35-86, get_xy_size :479-501,
                           finish_voxel_geometry :453-477, read_scans :394-451
  Nikon XTek (cone beam)   setup_experimental_geometry(numpy)
                           Core/src/Readers/xtek.cpp:57-129, get_xy_size
                           :760-798, finish_voxel_geometry :367-380,
                           read_scans :519-564, find_centre :671-758
  v padding                calc_v_alignment, Core/src/instruments.cpp:35-63
  block sizes              calculate_block_sizes, Core/src/voxels.cpp:12-46
Geometry is host-side (double); the scan transpose / -log / padding run on
the device with torch.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .projectors import Geometry

# aligned_allocator alignment without -mavx (Core/src/allocator.hpp:23-31):
# 16 bytes = 4 floats
V_ALIGN = 4


def calc_v_alignment(n: int, ppv: int, cone: bool):
    """(padded row count, data row offset) (instruments.cpp:35-63)."""
    nvox = n // ppv + (1 if n % ppv else 0)
    blocks = nvox // V_ALIGN + (1 if nvox % V_ALIGN else 0)
    if cone and blocks % 2:
        blocks += 1
    npix = blocks * V_ALIGN * ppv
    return npix, (npix - n) // 2


@dataclass
class Setup:
    """Result of an instrument set-up: projector geometry plus how the
    numpy scan maps into the padded sinogram."""
    geometry: Geometry
    data_v_size: int
    data_v_offset: int
    mask_radius: float = 0.0


def _fail(msg: str):
    raise ValueError(msg)


# ---- Diamond parallel beam -------------------------------------------------

def diamond_detector(n_angles: int, nv_data: int, nh: int, angles_deg,
                     rotation_centre: float, ppv: int):
    """h/v pixel positions and angles (diamond.cpp:35-86)."""
    ang = np.asarray(angles_deg, dtype=np.float32)
    if ang.size < 1:
        _fail("Bad angle array")
    if n_angles != ang.size:
        _fail("Number of projections doesn't match angle array")
    if nh < 1 or nv_data < 1:
        _fail("Bad array index for pixels")
    hsize = vsize = 1.0
    nv, voff = calc_v_alignment(nv_data, ppv, False)
    shift = (rotation_centre - nh / 2.0) * hsize if rotation_centre > 0.0 else 0.0
    h0 = -((nh - 1) * hsize) / 2.0 - shift
    h = h0 + np.arange(nh) * hsize
    v0 = -((nv - 1) * vsize) / 2.0
    v = v0 + np.arange(nv) * vsize
    rad = math.pi * (ang.astype(np.float64) / 180.0)
    return h, v, rad, nv, voff


def diamond_xy_size(h: np.ndarray, v: np.ndarray, ppv: int):
    """nx = ny, voxel sizes (diamond.cpp:479-501)."""
    step = h[1] - h[0]
    hmax = max(abs(h[0]), abs(h[-1]))
    n = int(math.ceil((2.0 * hmax) / step))
    nx = n // ppv + (1 if n % ppv else 0)
    hvox = step * ppv
    while (-hvox * nx / 2.0 >= h[0]) or (hvox * nx / 2.0 <= h[-1]):
        nx += 1
    vvox = (v[1] - v[0]) * ppv if v.size > 1 else ppv * 1.0
    return nx, nx, hvox, vvox


def diamond_setup(pix_shape, angles_deg, rotation_centre: float, ppv: int,
                  shape=None) -> Setup:
    """Full Diamond set-up from a [angle][v][h] scan shape; `shape` overrides
    the derived volume (pb_forward_project, diamond_wrapper.cpp:486-487)."""
    na, nv_data, nh = (int(q) for q in pix_shape)
    h, v, rad, nv, voff = diamond_detector(na, nv_data, nh, angles_deg, rotation_centre, ppv)
    nx, ny, hvox, vvox = diamond_xy_size(h, v, ppv)
    nz = nv // ppv + (1 if nv % ppv else 0)
    if shape is not None:
        nx, ny, nz = (int(q) for q in shape)
    vs = (hvox, hvox, vvox)
    origin = (-vs[0] * nx / 2.0, -vs[1] * ny / 2.0, -vs[2] * nz / 2.0)
    g = Geometry(rad, h, v, origin, vs, (nx, ny, nz))
    return Setup(g, nv_data, voff)


def diamond_read_scans(pixels, setup: Setup, is_log: bool):
    """[angle][v][h] scan -> padded [angle][h][v] sinogram, -log unless the
    data are already logged (diamond.cpp:394-451); device tensor in/out."""
    import torch
    na, nh, nv = setup.geometry.sino_shape
    out = torch.zeros((na, nh, nv), dtype=torch.float32, device=pixels.device)
    data = pixels.to(torch.float32).transpose(1, 2)
    if not is_log:
        data = -torch.log(data)
    out[:, :, setup.data_v_offset:setup.data_v_offset + setup.data_v_size] = data
    return out.contiguous()


# ---- Nikon XTek cone beam --------------------------------------------------

def xtek_setup(pix_shape, angles_deg, ppv: int, source_x: float, detector_x: float,
               pixel_h_size: float, pixel_v_size: float, mask_radius: float) -> Setup:
    """XTek set-up from a [angle][h][v] scan shape (xtek.cpp:57-129) plus the
    volume derived by get_xy_size / finish_voxel_geometry
    (xtek.cpp:760-798, 367-380)."""
    ang = np.asarray(angles_deg, dtype=np.float32)
    na, nh, nv_data = (int(q) for q in pix_shape)
    if ang.size < 1:
        _fail("Bad angle array")
    if na != ang.size:
        _fail("Number of projections doesn't match angle array")
    if nh < 1 or nv_data < 1:
        _fail("Bad array index for pixels")
    nv, voff = calc_v_alignment(nv_data, ppv, False)
    h = -((nh - 1) * pixel_h_size) / 2.0 + np.arange(nh) * pixel_h_size
    v = -((nv - 1) * pixel_v_size) / 2.0 + np.arange(nv) * pixel_v_size
    rad = math.pi * (ang.astype(np.float64) / 180.0)
    # get_xy_size
    n = nh // ppv + (1 if nh % ppv else 0)
    vsize = 2.0 * mask_radius / n
    vox_x = vsize * n / 2.0
    nz = nv // ppv + (1 if nv % ppv else 0)
    max_z = vsize * nz / 2.0
    v_pos = v[nv - 1] * (vox_x - source_x) / (detector_x - source_x)
    if max_z < v_pos:
        vsize = -2.0 * source_x * v[nv - 1] / (nz * (detector_x - source_x) - n * v[nv - 1])
        vsize *= 1.0 + 1.0 / nv
    vs = (vsize, vsize, vsize)
    origin = (-vs[0] * n / 2.0, -vs[1] * n / 2.0, -vs[2] * nz / 2.0)
    g = Geometry(rad, h, v, origin, vs, (n, n, nz), cone=True,
                 source=(source_x, 0.0, 0.0), detector_x=detector_x)
    return Setup(g, nv_data, voff, mask_radius)


def xtek_read_scans(pixels, setup: Setup, is_log: bool):
    """[angle][h][v] scan -> padded sinogram, -log unless logged
    (xtek.cpp:519-562).  find_centre is applied separately."""
    import torch
    na, nh, nv = setup.geometry.sino_shape
    out = torch.zeros((na, nh, nv), dtype=torch.float32, device=pixels.device)
    data = pixels.to(torch.float32)
    if not is_log:
        data = -torch.log(data)
    out[:, :, setup.data_v_offset:setup.data_v_offset + setup.data_v_size] = data
    return out.contiguous()


def _interp2d(h, ang, data, new_h, new_a):
    """Vectorised interpolate2D (xtek.cpp:566-669): data [na][nh] at
    positions (new_h, new_a); angles wrap at 2 pi."""
    nh, na = h.size, ang.size
    ph1 = np.minimum(np.searchsorted(h[1:], new_h, side="right"), nh - 1)
    exact_h = h[ph1] == new_h
    ph2 = np.minimum(ph1 + 1, nh - 1)
    two_pi = 2.0 * math.pi
    below = new_a < ang[0]
    pa1 = np.minimum(np.searchsorted(ang[1:], new_a, side="right"), na - 1)
    exact_a = (~below) & (ang[pa1] == new_a)
    last = (~below) & (pa1 == na - 1)
    a_lo = np.where(below, na - 1, pa1)
    a_hi = np.where(below | last, 0, np.minimum(pa1 + 1, na - 1))
    x1 = np.where(below, ang[na - 1] - two_pi, ang[a_lo])
    x2 = np.where(below, ang[0], np.where(last, ang[0] + two_pi, ang[a_hi]))
    f_lo_1 = data[a_lo, ph1].astype(np.float64)
    f_hi_1 = data[a_hi, ph1].astype(np.float64)
    f_lo_2 = data[a_lo, ph2].astype(np.float64)
    f_hi_2 = data[a_hi, ph2].astype(np.float64)
    hx1, hx2 = h[ph1], h[ph2]
    with np.errstate(divide="ignore", invalid="ignore"):
        lin_a = (f_lo_1 * (x2 - new_a) + f_hi_1 * (new_a - x1)) / (x2 - x1)
        lin_h = (f_lo_1 * (hx2 - new_h) + f_lo_2 * (new_h - hx1)) / (hx2 - hx1)
        bil = (f_lo_1 * (hx2 - new_h) * (x2 - new_a) + f_lo_2 * (new_h - hx1) * (x2 - new_a)
               + f_hi_1 * (hx2 - new_h) * (new_a - x1)
               + f_hi_2 * (new_h - hx1) * (new_a - x1)) / ((hx2 - hx1) * (x2 - x1))
    out = np.where(exact_h, np.where(exact_a, f_lo_1, lin_a),
                   np.where(exact_a, lin_h, bil))
    return out.astype(np.float32)


def find_centre(setup: Setup, sino_slice: np.ndarray) -> float:
    """Centre-of-rotation search of xtek.cpp:671-758 (T. Liu, Opt. Eng. 2009,
    046501) on one detector row ``sino_slice`` [angle][h]; shifts the
    geometry's source y and h pixels in place and returns the shift (mm)."""
    g = setup.geometry
    h, ph = g.h_pixels, g.angles
    na = ph.size
    sx = g.source[0]
    distance = g.detector_x - sx
    data = np.asarray(sino_slice, np.float32)
    midpoint = 0.0
    nsamples = 21
    gamma_i = np.arctan(h / distance)
    for prec in (1.0, 0.1, 0.01, 0.001, 0.0001):
        scor = midpoint - (nsamples // 2) * prec
        M = np.zeros(nsamples, np.float32)
        for j in range(nsamples):
            gamma_c = math.atan((midpoint + (j - nsamples // 2) * prec) / distance)
            beta = math.pi + 2.0 * (gamma_i - gamma_c)
            s2 = distance * np.tan(2.0 * gamma_c - gamma_i)
            ks = np.nonzero((s2 >= h[0]) & (s2 <= h[-1]))[0]
            if ks.size == 0:
                M[j] = np.float32(np.nan)
                continue
            ab = ph[:, None] + beta[None, ks]
            # wrap into [ph[0], ph[na-1]] like the reference's two loops
            two_pi = 2.0 * math.pi
            lo = ab < ph[0]
            ab = np.where(lo, ab + two_pi * np.ceil((ph[0] - ab) / two_pi), ab)
            hi = ab > ph[na - 1]
            ab = np.where(hi, ab - two_pi * np.ceil((ab - ph[na - 1]) / two_pi), ab)
            p = _interp2d(h, ph, data, np.broadcast_to(s2[ks], ab.shape), ab)
            t = data[:, ks] - p
            M[j] = np.float32(np.sum(t * t, dtype=np.float32) / np.float32(t.size))
        valid = np.where(np.isnan(M), np.inf, M)
        ind = int(np.argmin(valid))
        if not np.isfinite(valid[ind]):
            _fail("Failure in XTek find centre")
        midpoint = scor + ind * prec
    y_centre = midpoint * sx / distance
    g.source = (g.source[0], g.source[1] + y_centre, g.source[2])
    g.h_pixels = g.h_pixels + y_centre
    g._c = None
    return y_centre
