"""``ccpi.reconstruction.parallelbeam``: drop-in for the reference module of
the same name (Wrappers/Python/src/diamond_module.cpp:23-56,
diamond_wrapper.cpp).  ``alg`` holds the functions; they take numpy arrays
and return numpy arrays like the Boost.Python wrappers, and run on the
current HIP device.

Behaviour kept from the reference path (voxels.cpp:131-205):
  * input scans are [angle][v][h] float32, angles in degrees;
  * Diamond geometry: 1 mm pixels, rotation-centre shift, v padded to the
    alignment, voxels derived from the detector (diamond.cpp:35-86,
    479-501);
  * -log unless is_pixels_in_log; the result is clamped at 0 (voxels.cpp:187);
  * nthreads is accepted and ignored (no OpenMP here).
Differences: invalid input raises instead of returning a (1,1,1) zero array
(diamond_wrapper.cpp:111-127); CGLS's residual norms are returned by
``alg.cgls_norms`` (the reference's norm output is never filled,
diamond_wrapper.cpp:99-108).
"""

from __future__ import annotations

import types

import numpy as np

from . import instruments as _ins
from . import projectors as _pj

alg = types.ModuleType("ccpi.reconstruction.parallelbeam.alg")
_last = {}


def _device():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("ccpi.reconstruction needs a HIP device")
    return torch.device("cuda", torch.cuda.current_device())


def _as_f32(a, name):
    a = np.asarray(a)
    if a.dtype != np.float32:
        a = a.astype(np.float32)
    return np.ascontiguousarray(a)


def _norm_quirk(norm_r, niterations):
    """The reference never writes residual norms into ``norm_r``: when its
    first element is nonzero it overwrites the first ``niterations``
    entries with zeros, otherwise it leaves the array alone
    (diamond_wrapper.cpp:95-104, 213-222).  Kept as is; the norms are
    available from ``alg.cgls_norms()``."""
    if isinstance(norm_r, np.ndarray) and norm_r.size and norm_r.flat[0] != 0:
        norm_r.reshape(-1)[:int(niterations)] = 0


def _reconstruct(pixels, angles, rotation_centre, resolution, niterations, is_log, algo,
                 x0=None, lam=0.0):
    import torch
    dev = _device()
    pix = _as_f32(pixels, "pixels")
    if pix.ndim != 3:
        raise ValueError("pixels must be [angle][v][h]")
    setup = _ins.diamond_setup(pix.shape, angles, float(rotation_centre), int(resolution))
    g = setup.geometry
    b = _ins.diamond_read_scans(torch.from_numpy(pix).to(dev), setup, bool(is_log))
    x = torch.zeros(g.shape, dtype=torch.float32, device=dev)
    if x0 is not None:
        x0 = _as_f32(x0, "last_iteration_volume")
        if x0.shape != g.shape:
            raise ValueError(f"previous volume must be {g.shape}")
        x.copy_(torch.from_numpy(x0))
    if algo == "cgls":
        x, _, norms = _pj.cgls(g, b, int(niterations), x)
        _last["norms"] = norms
    elif algo in ("tikhonov", "tv"):
        kind = _pj.TIKHONOV if algo == "tikhonov" else _pj.TV
        x, _, norms = _pj.cgls_reg(g, b, int(niterations), kind, float(lam), x)
        _last["norms"] = norms
    elif algo == "sirt":
        x = _pj.sirt(g, b, int(niterations), x)
    elif algo == "mlem":
        x = _pj.mlem(g, b, int(niterations), x)
    x.clamp_(min=0.0)
    return x.cpu().numpy()


def cgls(pixels, angles, rotation_centre, resolution, niterations, nthreads,
         is_pixels_in_log=True):
    """CGLS (diamond_wrapper.cpp:259-266)."""
    return _reconstruct(pixels, angles, rotation_centre, resolution, niterations,
                        is_pixels_in_log, "cgls")


def sirt(pixels, angles, rotation_centre, resolution, niterations, nthreads,
         is_pixels_in_log=True):
    """SIRT (diamond_wrapper.cpp:282-289)."""
    return _reconstruct(pixels, angles, rotation_centre, resolution, niterations,
                        is_pixels_in_log, "sirt")


def mlem(pixels, angles, rotation_centre, resolution, niterations, nthreads,
         is_pixels_in_log=True):
    """MLEM (diamond_wrapper.cpp:305-312)."""
    return _reconstruct(pixels, angles, rotation_centre, resolution, niterations,
                        is_pixels_in_log, "mlem")


def cgls_step(pixels, angles, rotation_centre, resolution, niterations, nthreads,
              is_pixels_in_log, last_iteration_volume):
    """CGLS from a previous volume: x0 + CGLS(b), the residual is not formed
    from x0 (cgls.cpp:45-49, voxels.cpp:295-371)."""
    return _reconstruct(pixels, angles, rotation_centre, resolution, niterations,
                        is_pixels_in_log, "cgls", last_iteration_volume)


def sirt_step(pixels, angles, rotation_centre, resolution, niterations, nthreads,
              is_pixels_in_log, last_iteration_volume):
    """SIRT continued from a previous volume."""
    return _reconstruct(pixels, angles, rotation_centre, resolution, niterations,
                        is_pixels_in_log, "sirt", last_iteration_volume)


def mlem_step(pixels, angles, rotation_centre, resolution, niterations, nthreads,
              is_pixels_in_log, last_iteration_volume):
    """MLEM ignores the previous volume: it restarts from ones (mlem.cpp:30)."""
    return _reconstruct(pixels, angles, rotation_centre, resolution, niterations,
                        is_pixels_in_log, "mlem", last_iteration_volume)


def cgls_conv(pixels, angles, rotation_centre, resolution, niterations, nthreads, norm_r,
              is_pixels_in_log=True):
    """CGLS with a norm_r argument (diamond_wrapper.cpp:388-396)."""
    out = _reconstruct(pixels, angles, rotation_centre, resolution, niterations,
                       is_pixels_in_log, "cgls")
    _norm_quirk(norm_r, niterations)
    return out


def cgls_conv_step(pixels, angles, rotation_centre, resolution, niterations, nthreads, norm_r,
                   is_pixels_in_log, last_iteration_volume):
    """cgls_conv from a previous volume (diamond_wrapper.cpp:398-409)."""
    out = _reconstruct(pixels, angles, rotation_centre, resolution, niterations,
                       is_pixels_in_log, "cgls", last_iteration_volume)
    _norm_quirk(norm_r, niterations)
    return out


def cgls_tikhonov(pixels, angles, rotation_centre, resolution, niterations, nthreads,
                  regularize, norm_r, is_pixels_in_log=True):
    """CGLS with a Tikhonov term (diamond_wrapper.cpp:329-339, cgls.cpp:
    343-430)."""
    out = _reconstruct(pixels, angles, rotation_centre, resolution, niterations,
                       is_pixels_in_log, "tikhonov", lam=regularize)
    _norm_quirk(norm_r, niterations)
    return out


def cgls_tikhonov_step(pixels, angles, rotation_centre, resolution, niterations, nthreads,
                       regularize, norm_r, is_pixels_in_log, last_iteration_volume):
    """cgls_tikhonov from a previous volume (diamond_wrapper.cpp:341-352)."""
    out = _reconstruct(pixels, angles, rotation_centre, resolution, niterations,
                       is_pixels_in_log, "tikhonov", last_iteration_volume, regularize)
    _norm_quirk(norm_r, niterations)
    return out


def cgls_TVreg(pixels, angles, rotation_centre, resolution, niterations, nthreads, regularize,
               norm_r, is_pixels_in_log=True):
    """CGLS with a TV term (diamond_wrapper.cpp:358-368, cgls.cpp:343-435)."""
    out = _reconstruct(pixels, angles, rotation_centre, resolution, niterations,
                       is_pixels_in_log, "tv", lam=regularize)
    _norm_quirk(norm_r, niterations)
    return out


def cgls_TVreg_step(pixels, angles, rotation_centre, resolution, niterations, nthreads,
                    regularize, norm_r, is_pixels_in_log, last_iteration_volume):
    """cgls_TVreg from a previous volume (diamond_wrapper.cpp:370-381)."""
    out = _reconstruct(pixels, angles, rotation_centre, resolution, niterations,
                       is_pixels_in_log, "tv", last_iteration_volume, regularize)
    _norm_quirk(norm_r, niterations)
    return out


def cgls_norms():
    """||A^T r|| per iteration of the last CGLS-family call."""
    return np.array(_last.get("norms", np.zeros(0)))


def pb_setup_geometry_from_image(volume, angles, resolution):
    """Geometry derived from a volume shape (diamond_wrapper.cpp:558-631)."""
    vol = np.asarray(volume)
    width = max(vol.shape[0], vol.shape[1])
    setup = _ins.diamond_setup((np.asarray(angles).size, vol.shape[2], width), angles,
                               width / 2.0, int(resolution))
    na, nh, nv = setup.geometry.sino_shape
    return {"n_angles": na, "n_h": nh, "n_v": nv, "output_volume_x": vol.shape[0],
            "output_volume_y": vol.shape[1], "output_volume_z": vol.shape[2]}


def pb_setup_geometry_from_acquisition(projections, angles, rotation_centre, resolution):
    """Geometry derived from a [angle][v][h] stack."""
    p = np.asarray(projections)
    setup = _ins.diamond_setup(p.shape, angles, float(rotation_centre), int(resolution))
    na, nh, nv = setup.geometry.sino_shape
    nx, ny, nz = setup.geometry.shape
    return {"n_angles": na, "n_h": nh, "n_v": nv, "output_volume_x": nx,
            "output_volume_y": ny, "output_volume_z": nz}


def pb_forward_project(volume, angles, resolution):
    """Forward projection of a [x][y][z] volume; the detector is derived from
    the volume shape: width max(nx, ny), centre width / 2, height nz
    (diamond_wrapper.cpp:421-556).  Returns [angle][v][h]."""
    import torch
    dev = _device()
    vol = _as_f32(volume, "volume")
    width = max(vol.shape[0], vol.shape[1])
    setup = _ins.diamond_setup((np.asarray(angles).size, vol.shape[2], width), angles,
                               width / 2.0, int(resolution), shape=vol.shape)
    g = setup.geometry
    y = _pj.forward_project(g, torch.from_numpy(vol).to(dev))
    return y.transpose(1, 2).contiguous().cpu().numpy()


def pb_backward_project(projections, angles, rotation_centre, resolution):
    """Backprojection of a [angle][v][h] stack onto the volume derived from
    it (diamond_wrapper.cpp:633-757).  Rows beyond the stack read as 0 (the
    reference reads past the array there)."""
    import torch
    dev = _device()
    p = _as_f32(projections, "projections")
    setup = _ins.diamond_setup(p.shape, angles, float(rotation_centre), int(resolution))
    g = setup.geometry
    na, nh, nv = g.sino_shape
    s = torch.zeros((na, nh, nv), dtype=torch.float32, device=dev)
    rows = min(nv, p.shape[1])
    s[:, :, :rows] = torch.from_numpy(p[:, :rows, :]).to(dev).transpose(1, 2)
    return _pj.backward_project(g, s.contiguous()).cpu().numpy()


for _f in (cgls, sirt, mlem, cgls_step, sirt_step, mlem_step, cgls_conv, cgls_conv_step,
           cgls_tikhonov, cgls_tikhonov_step, cgls_TVreg, cgls_TVreg_step, cgls_norms,
           pb_setup_geometry_from_image, pb_setup_geometry_from_acquisition,
           pb_forward_project, pb_backward_project):
    setattr(alg, _f.__name__, _f)
