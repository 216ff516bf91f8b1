"""The cone-beam Python drop-in end to end (GPU tests): a Nikon XTek scan
written as files (.xtekct, _ctdata.txt, 16-bit TIFFs), read by
conebeam.alg.load_xtek and reconstructed by conebeam.alg.cgls /
cgls_conv / cgls_tikhonov / cgls_TVreg, against the oracle's CGLS on the
geometry the call reports (after its centre search) and the same ingested
sinogram, with the final clamp."""

import math
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
PIL = pytest.importorskip("PIL")
from PIL import Image

import oracle
from geometries import ellipsoid_phantom

pytestmark = pytest.mark.gpu

NA, NH, NV = 36, 24, 16
SOD, SDD, PIX, MASK = 60.0, 120.0, 0.2, 1.1


def _og(g):
    return oracle.Geometry(g.angles, g.h_pixels, g.v_pixels, g.origin, g.voxel_size, g.shape,
                           cone=True, source=g.source, det_x=g.detector_x)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.fixture(scope="module")
def scan(tmp_path_factory):
    """An XTek scan of a 5-ellipsoid phantom: intensities I0 exp(-p) as
    16-bit TIFFs (row 0 at the top), I0 = 60000 = WhiteLevel."""
    from ccpi.reconstruction import instruments as ins
    d = tmp_path_factory.mktemp("xtek")
    ang = np.linspace(0, 360, NA, endpoint=False)
    setup = ins.xtek_setup((NA, NH, NV), ang.astype(np.float32), 1, -SOD, SDD - SOD, PIX, PIX,
                           MASK)
    g = setup.geometry
    ph = ellipsoid_phantom(g.shape, g.voxel_size, g.origin, 5, seed=9, dens=(0.5, 2.0))
    p = _og(g).fp(ph)
    counts = np.clip(np.round(60000.0 * np.exp(-p)), 0, 65535).astype(np.uint16)
    cfg = ["[XTekCT]", "Name=ph", f"SrcToObject={SOD}", f"SrcToDetector={SDD}",
           f"MaskRadius={MASK}", f"DetectorPixelsX={NH}", f"DetectorPixelsY={NV}",
           f"DetectorPixelSizeX={PIX}", f"DetectorPixelSizeY={PIX}", f"Projections={NA}",
           "InitialAngle=0", "WhiteLevel=60000", "Scattering=0"]
    (d / "ph.xtekct").write_text("\n".join(cfg) + "\n")
    with open(d / "_ctdata.txt", "w") as f:
        f.write("P A T\n0 0 0\nP A T\n")
        for i, a in enumerate(ang):
            f.write(f"{i + 1} {a:.6f} 0\n")
    for i in range(NA):
        img = counts[i].T[::-1]          # [v][h], top row = last detector row
        Image.frombytes("I;16", (NH, NV), np.ascontiguousarray(img).astype("<u2").tobytes()
                        ).save(d / f"ph{i + 1:04d}.tif")
    return str(d / "ph.xtekct"), p


def _call(alg, name, pix, ang, sx, dx, ph, pv, mr, niter, **kw):
    args = (pix, ang, np.zeros(1, np.float32), np.zeros(1, np.float32), 1, sx, dx, ph, pv, mr,
            False, np.zeros(3, np.float32), np.ones(3, np.float32), niter, 1)
    if name == "cgls":
        return alg.cgls(*args, True)
    if name == "cgls_conv":
        return alg.cgls_conv(*args, np.zeros(niter, np.float32), True)
    return getattr(alg, name)(*args, kw["lam"], np.zeros(niter, np.float32), True)


@pytest.mark.parametrize("name", ["cgls", "cgls_conv", "cgls_tikhonov", "cgls_TVreg"])
def test_load_xtek_then_reconstruct(scan, name):
    from ccpi.reconstruction.conebeam import alg
    path, p = scan
    pix, ang, sx, dx, ph, pv, mr = alg.load_xtek(path)
    assert pix.shape == (NA, NH, NV) and (sx, dx, mr) == (SOD, SDD, MASK)
    # -log of the quantised intensities gives back the line integrals
    assert np.abs(pix - p).max() < 1e-3
    print(f"\nxtek {name}", flush=True)
    lam = 0.05 if name == "cgls_tikhonov" else 1e-4
    vox = _call(alg, name, pix, ang, sx, dx, ph, pv, mr, 4, lam=lam)
    og = _og(alg.last_geometry())
    if name in ("cgls", "cgls_conv"):
        xo, _, _ = oracle.cgls(og, pix, 4)
    else:
        kind = oracle.TIKHONOV if name == "cgls_tikhonov" else oracle.TV
        xo, _, _ = oracle.cgls_reg(og, pix, 4, kind, lam)
    assert _rel(vox, np.maximum(xo, 0)) < 2e-4
