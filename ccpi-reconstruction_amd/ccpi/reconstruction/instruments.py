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
the device with torch, and find_centre's search runs in the HIP library.
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


def find_centre(setup: Setup, sino, v_slice: int) -> float:
    """Centre-of-rotation search of xtek.cpp:671-758 (T. Liu, Opt. Eng. 2009,
    046501) on row ``v_slice`` of the device sinogram ``sino``
    [angle][h][v], run by the HIP library (ccpi_xtek_find_centre).  Shifts
    the geometry's source y and h pixels in place (xtek.cpp:755-756) and
    returns the shift (mm)."""
    import ctypes as C
    from ._lib import check, lib
    from .projectors import _dev, _stream
    g = setup.geometry
    y = C.c_double()
    check(lib().ccpi_xtek_find_centre(C.byref(g.c_struct()), _dev(sino, g.sino_shape, "sino"),
                                      int(v_slice), C.byref(y), _stream()))
    y_centre = y.value
    g.source = (g.source[0], g.source[1] + y_centre, g.source[2])
    g.h_pixels = g.h_pixels + y_centre
    g._c = None
    return y_centre
