"""``ccpi.reconstruction.conebeam``: drop-in for the reference module of the
same name (Wrappers/Python/src/conebeam_module.cpp:23-42,
conebeam_wrapper.cpp).

Behaviour kept from the reference path (conebeam_wrapper.cpp:14-121,
voxels.cpp:208-290):
  * scans are [angle][h][v] float32, angles in degrees;
  * source_x = -source_to_sample, detector_x = source_to_detector +
    source_x; XTek geometry and voxel size from the mask radius
    (xtek.cpp:57-129, 760-798);
  * -log unless is_pixels_in_log, then the find_centre search shifts the
    source and h pixels (xtek.cpp:563, 671-758);
  * the result is clamped at 0; nthreads is ignored.
The reference overwrites its full_vox_origin / voxel_size arguments in C++
only; here they are returned by ``alg.last_geometry()``.
``create_sinogram`` is the static cone f2D with explicit geometry and angles
in radians (conebeam_wrapper.cpp:346-400).
"""

from __future__ import annotations

import types

import numpy as np

from . import instruments as _ins
from . import projectors as _pj

alg = types.ModuleType("ccpi.reconstruction.conebeam.alg")
_last = {}


def _device():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("ccpi.reconstruction needs a HIP device")
    return torch.device("cuda", torch.cuda.current_device())


def _reconstruct(algo, pixels, angles, h_offsets, v_offsets, pixels_per_voxel,
                 source_to_sample, source_to_detector, pixel_h_size, pixel_v_size,
                 mask_radius, beam_harden, full_vox_origin, voxel_size, niterations,
                 nthreads, is_pixels_in_log, lam=0.0):
    import torch
    dev = _device()
    pix = np.ascontiguousarray(np.asarray(pixels, dtype=np.float32))
    if pix.ndim != 3:
        raise ValueError("pixels must be [angle][h][v]")
    sx = -float(source_to_sample)
    dx = float(source_to_detector) + sx
    setup = _ins.xtek_setup(pix.shape, angles, int(pixels_per_voxel), sx, dx,
                            float(pixel_h_size), float(pixel_v_size), float(mask_radius))
    b = _ins.xtek_read_scans(torch.from_numpy(pix).to(dev), setup, bool(is_pixels_in_log))
    g = setup.geometry
    _ins.find_centre(setup, b, g.v_pixels.size // 2)
    x = torch.zeros(g.shape, dtype=torch.float32, device=dev)
    if algo == "cgls":
        x, _, norms = _pj.cgls(g, b, int(niterations), x)
        _last["norms"] = norms
    elif algo in ("tikhonov", "tv"):
        kind = _pj.TIKHONOV if algo == "tikhonov" else _pj.TV
        x, _, norms = _pj.cgls_reg(g, b, int(niterations), kind, float(lam), x)
        _last["norms"] = norms
    elif algo == "sirt":
        x = _pj.sirt(g, b, int(niterations), x)
    else:
        x = _pj.mlem(g, b, int(niterations), x)
    x.clamp_(min=0.0)
    _last["geometry"] = g
    return x.cpu().numpy()


def cgls(pixels, angles, h_offsets, v_offsets, pixels_per_voxel, source_to_sample,
         source_to_detector, pixel_h_size, pixel_v_size, mask_radius, beam_harden,
         full_vox_origin, voxel_size, niterations, nthreads, is_pixels_in_log=True):
    """CGLS (conebeam_wrapper.cpp:124-138)."""
    return _reconstruct("cgls", pixels, angles, h_offsets, v_offsets, pixels_per_voxel,
                        source_to_sample, source_to_detector, pixel_h_size, pixel_v_size,
                        mask_radius, beam_harden, full_vox_origin, voxel_size, niterations,
                        nthreads, is_pixels_in_log)


def sirt(pixels, angles, h_offsets, v_offsets, pixels_per_voxel, source_to_sample,
         source_to_detector, pixel_h_size, pixel_v_size, mask_radius, beam_harden,
         full_vox_origin, voxel_size, niterations, nthreads, is_pixels_in_log=True):
    """SIRT (conebeam_wrapper.cpp:140-154)."""
    return _reconstruct("sirt", pixels, angles, h_offsets, v_offsets, pixels_per_voxel,
                        source_to_sample, source_to_detector, pixel_h_size, pixel_v_size,
                        mask_radius, beam_harden, full_vox_origin, voxel_size, niterations,
                        nthreads, is_pixels_in_log)


def mlem(pixels, angles, h_offsets, v_offsets, pixels_per_voxel, source_to_sample,
         source_to_detector, pixel_h_size, pixel_v_size, mask_radius, beam_harden,
         full_vox_origin, voxel_size, niterations, nthreads, is_pixels_in_log=True):
    """MLEM (conebeam_wrapper.cpp:156-170)."""
    return _reconstruct("mlem", pixels, angles, h_offsets, v_offsets, pixels_per_voxel,
                        source_to_sample, source_to_detector, pixel_h_size, pixel_v_size,
                        mask_radius, beam_harden, full_vox_origin, voxel_size, niterations,
                        nthreads, is_pixels_in_log)


def _norm_quirk(norm_r, niterations):
    """The reference zeroes norm_r[:niterations] when norm_r tests nonzero
    and otherwise leaves it alone; it never stores the residual norms
    (conebeam_wrapper.cpp:84-93).  A norm_r longer than one element makes
    the reference's truth test raise; here its first element decides."""
    if isinstance(norm_r, np.ndarray) and norm_r.size and norm_r.flat[0] != 0:
        norm_r.reshape(-1)[:int(niterations)] = 0


def cgls_conv(pixels, angles, h_offsets, v_offsets, pixels_per_voxel, source_to_sample,
              source_to_detector, pixel_h_size, pixel_v_size, mask_radius, beam_harden,
              full_vox_origin, voxel_size, niterations, nthreads, norm_r,
              is_pixels_in_log=True):
    """CGLS with a norm_r argument (conebeam_wrapper.cpp:215-231)."""
    out = _reconstruct("cgls", pixels, angles, h_offsets, v_offsets, pixels_per_voxel,
                       source_to_sample, source_to_detector, pixel_h_size, pixel_v_size,
                       mask_radius, beam_harden, full_vox_origin, voxel_size, niterations,
                       nthreads, is_pixels_in_log)
    _norm_quirk(norm_r, niterations)
    return out


def cgls_tikhonov(pixels, angles, h_offsets, v_offsets, pixels_per_voxel, source_to_sample,
                  source_to_detector, pixel_h_size, pixel_v_size, mask_radius, beam_harden,
                  full_vox_origin, voxel_size, niterations, nthreads, regularize, norm_r,
                  is_pixels_in_log=True):
    """CGLS with a Tikhonov term (conebeam_wrapper.cpp:172-189)."""
    out = _reconstruct("tikhonov", pixels, angles, h_offsets, v_offsets, pixels_per_voxel,
                       source_to_sample, source_to_detector, pixel_h_size, pixel_v_size,
                       mask_radius, beam_harden, full_vox_origin, voxel_size, niterations,
                       nthreads, is_pixels_in_log, regularize)
    _norm_quirk(norm_r, niterations)
    return out


def cgls_TVreg(pixels, angles, h_offsets, v_offsets, pixels_per_voxel, source_to_sample,
               source_to_detector, pixel_h_size, pixel_v_size, mask_radius, beam_harden,
               full_vox_origin, voxel_size, niterations, nthreads, regularize, norm_r,
               is_pixels_in_log=True):
    """CGLS with a TV term (conebeam_wrapper.cpp:191-213)."""
    out = _reconstruct("tv", pixels, angles, h_offsets, v_offsets, pixels_per_voxel,
                       source_to_sample, source_to_detector, pixel_h_size, pixel_v_size,
                       mask_radius, beam_harden, full_vox_origin, voxel_size, niterations,
                       nthreads, is_pixels_in_log, regularize)
    _norm_quirk(norm_r, niterations)
    return out


# XTek phantom scan (Readers/xtek.cpp:131-173, 392-457)
PHANTOM_N = 512
PHANTOM_PIXEL = 0.390625
PHANTOM_CUBES = (((107, 189), (107, 189), (57, 139)),
                 ((189, 271), (189, 271), (139, 221)),
                 ((271, 353), (271, 353), (221, 303)))


def phantom_setup():
    """The phantom scan's geometry (xtek.cpp:131-173): source at x = -250,
    detector at x = 737, 512 x 512 pixels of 0.390625 mm centred on the
    axis, 250 angles over [0, 2 pi), mask radius from the outermost pixel.
    Returns (geometry of the 512^3 phantom volume, angles in radians,
    mask radius); the volume's voxel size is 2 mask_radius / 512 and it is
    centred on the origin (xtek.cpp:392-414)."""
    src_x, det_x = -250.0, 737.0
    first = -99.8046875
    c = 0
    while c < 1000 and first + c * PHANTOM_PIXEL < 100.0 + 0.001:
        c += 1
    pix = first + np.arange(c, dtype=np.float64) * PHANTOM_PIXEL
    na = 250
    phi = np.arange(na, dtype=np.float64) * (2.0 * np.pi / na)
    mask = -src_x * np.sin(np.arctan(pix[-1] / (det_x - src_x)))
    n = PHANTOM_N
    vox = 2.0 * mask / n
    off = -(vox * n) / 2.0
    g = _pj.Geometry(phi, pix, pix.copy(), (off, off, off), (vox, vox, vox), (n, n, n),
                     cone=True, source=(src_x, 0.0, 0.0), detector_x=det_x)
    return g, phi, float(mask)


def _phantom_volume(dev):
    """512^3 zeros with three unit cubes (xtek.cpp:424-446)."""
    import torch
    n = PHANTOM_N
    x = torch.zeros((n, n, n), dtype=torch.float32, device=dev)
    for (i0, i1), (j0, j1), (k0, k1) in PHANTOM_CUBES:
        x[i0:i1, j0:j1, k0:k1] = 1.0
    return x


def _phantom_scalars(g, mask):
    # the wrapper reports source_to_sample / source_to_detector
    # (conebeam_wrapper.cpp:245-248)
    sx = -g.source[0]
    return sx, g.detector_x + sx, PHANTOM_PIXEL, PHANTOM_PIXEL, mask


def create_phantom():
    """Projections of the XTek cube phantom (conebeam_wrapper.cpp:238-268):
    (pixels [250][512][512] as [angle][h][v], angles in degrees (float32),
    source_to_sample 250.0, source_to_detector 987.0, pixel_h_size,
    pixel_v_size, mask_radius).  The projections come from this library's
    cone forward projector; the reference uses its generic 3-D Siddon
    (project_line.hpp) for this one call.  Both give exact ray-voxel
    intersection lengths."""
    g, phi, mask = phantom_setup()
    dev = _device()
    y = _pj.forward_project(g, _phantom_volume(dev))
    ang = (phi * 180.0 / np.pi).astype(np.float32)
    return (y.cpu().numpy(), ang) + _phantom_scalars(g, mask)


def create_phantom_volume():
    """The phantom volume itself (conebeam_wrapper.cpp:270-311): (voxels
    [512][512][512], angles in degrees, h_pixels, v_pixels (float32),
    source_to_sample, source_to_detector, pixel_h_size, pixel_v_size,
    mask_radius)."""
    g, phi, mask = phantom_setup()
    vol = _phantom_volume(_device()).cpu().numpy()
    ang = (phi * 180.0 / np.pi).astype(np.float32)
    return ((vol, ang, g.h_pixels.astype(np.float32), g.v_pixels.astype(np.float32))
            + _phantom_scalars(g, mask))


def load_xtek(filename):
    """Read a Nikon XTek scan from its .xtekct file
    (conebeam_wrapper.cpp:313-344): (pixels [angle][h][v] after -log,
    angles in degrees, source_to_sample, source_to_detector, 0.2, 0.2,
    mask_radius); see ``formats.load_xtek``."""
    from . import formats
    return formats.load_xtek(filename)


def last_geometry():
    """Geometry (with the find_centre shift) of the last reconstruction."""
    return _last.get("geometry")


def cgls_norms():
    return np.array(_last.get("norms", np.zeros(0)))


def create_sinogram(source_x, source_y, source_z, detector_x, h_pixels, v_pixels, angles,
                    voxels, grid_offset, voxel_size):
    """Static cone f2D with explicit geometry; angles in radians;
    returns [angle][h][v] (conebeam_wrapper.cpp:346-400)."""
    import torch
    dev = _device()
    vol = np.ascontiguousarray(np.asarray(voxels, dtype=np.float32))
    f = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)
    g = _pj.Geometry(f(angles), f(h_pixels), f(v_pixels), f(grid_offset)[:3],
                     f(voxel_size)[:3], vol.shape, cone=True,
                     source=(float(np.float32(source_x)), float(np.float32(source_y)),
                             float(np.float32(source_z))),
                     detector_x=float(np.float32(detector_x)))
    return _pj.forward_project(g, torch.from_numpy(vol).to(dev)).cpu().numpy()


for _f in (cgls, cgls_conv, sirt, mlem, cgls_tikhonov, cgls_TVreg, last_geometry, cgls_norms,
           create_sinogram, create_phantom, create_phantom_volume, load_xtek):
    setattr(alg, _f.__name__, _f)
