"""Scan readers and result writers (CPU tests).

TIFF: the reader is checked against files written by Pillow (an
independent TIFF implementation) in both byte orders, uncompressed and
with LZW, PackBits and Deflate strips (and LZW with the horizontal
predictor), its PackBits decoder against the TIFF 6.0 specification's
example, and the writer by reading its files back with Pillow.  The XTek reader runs on synthetic
scans written here (.xtekct, _ctdata.txt or .ang, 16-bit TIFFs from
Pillow): the layout ([angle][h][v], v flipped from image rows and centred
in the aligned rows), the angle lists (the .ang file in reverse), the
white-level / scattering normalisation and -log of xtek.cpp:477-512, and
load_xtek's returned scalars (conebeam_wrapper.cpp:313-344).  The result
writers follow results.cpp: TIFF slices by truncation (clamped or
range-stretched), the raw dump appended, the .bgs header and [z][y][x]
data.
"""

import math
import os

import numpy as np
import pytest

PIL = pytest.importorskip("PIL")
from PIL import Image

from ccpi.reconstruction import formats as F


def _pil16(path, a, big):
    mode, dt = ("I;16B", ">u2") if big else ("I;16", "<u2")
    Image.frombytes(mode, (a.shape[1], a.shape[0]), a.astype(dt).tobytes()).save(path)


@pytest.mark.parametrize("big", [False, True])
def test_read_tiff16_matches_pillow(tmp_path, big):
    a = (np.arange(33 * 19).reshape(19, 33) * 997 % 65536).astype(np.uint16)
    p = str(tmp_path / "a.tif")
    _pil16(p, a, big)
    assert np.array_equal(F.read_tiff16(p), a)


@pytest.mark.parametrize("comp", ["tiff_lzw", "packbits", "tiff_adobe_deflate"])
@pytest.mark.parametrize("big", [False, True])
def test_read_tiff16_compressed_strips(tmp_path, comp, big):
    """LZW (table clears and 12-bit codes on noise), PackBits and Deflate
    strips as Pillow's libtiff writes them, in both byte orders."""
    rng = np.random.default_rng(5)
    a = rng.integers(0, 65536, (257, 300)).astype(np.uint16)
    a[100:140, :] = 4321                     # long runs
    a[:, 7] = np.arange(257) * 255
    mode, dt = ("I;16B", ">u2") if big else ("I;16", "<u2")
    p = str(tmp_path / "c.tif")
    Image.frombytes(mode, (a.shape[1], a.shape[0]), a.astype(dt).tobytes()).save(
        p, compression=comp)
    assert np.array_equal(F.read_tiff16(p), a)


def test_read_tiff16_lzw_with_horizontal_predictor(tmp_path):
    a = (np.add.outer(np.arange(64) * 3, np.arange(80) * 700) % 65536).astype(np.uint16)
    p = str(tmp_path / "p.tif")
    Image.frombytes("I;16", (80, 64), a.astype("<u2").tobytes()).save(
        p, compression="tiff_lzw", tiffinfo={317: 2})
    with open(p, "rb") as f:
        _, tags = F._tiff_tags(f.read())
    assert tags[259] == (5,) and tags[317] == (2,)
    assert np.array_equal(F.read_tiff16(p), a)


def test_unpackbits_known_answer():
    """The PackBits example of the TIFF 6.0 specification (section 9)."""
    packed = bytes.fromhex("FE AA 02 80 00 2A FD AA 03 80 00 2A 22 F7 AA")
    want = bytes.fromhex("AA AA AA 80 00 2A AA AA AA AA 80 00 2A 22 AA AA AA AA AA AA AA AA AA AA")
    assert F._unpackbits(packed) == want


def test_read_tiff16_rejects_8bit(tmp_path):
    p = str(tmp_path / "b.tif")
    Image.fromarray(np.zeros((4, 5), np.uint8)).save(p)
    with pytest.raises(F.FormatError, match="16bit"):
        F.read_tiff16(p)


@pytest.mark.parametrize("dt", [np.uint8, np.uint16])
@pytest.mark.parametrize("rows", [1, 7])
def test_write_tiff_bottom_up_read_by_pillow(tmp_path, dt, rows):
    a = (np.arange(rows * 9).reshape(rows, 9) * 13).astype(dt)
    p = str(tmp_path / "w.tif")
    F.write_tiff(p, a)
    assert np.array_equal(np.array(Image.open(p)), a[::-1])


def _write_scan(tmp, na, nh, nvd, wl=4000.0, sc=2.5, init=10.0, use_ang=False, big=False,
                sep="_", seed=0):
    rng = np.random.default_rng(seed)
    imgs = rng.integers(0, 4200, size=(na, nvd, nh)).astype(np.uint16)
    imgs[0, 0, 0] = 0
    cfg = ["[XTekCT]", "Name=scan", "SrcToObject=120.5", "SrcToDetector=700.25",
           "MaskRadius=11.75", f"DetectorPixelsX={nh}", f"DetectorPixelsY={nvd}",
           "DetectorPixelSizeX=0.4", "DetectorPixelSizeY=0.5", f"Projections={na}",
           f"InitialAngle={init}", f"WhiteLevel={wl}", f"Scattering={sc}",
           f"InputSeparator={sep}", "CoefX2=1.0"]
    with open(os.path.join(tmp, "scan.xtekct"), "w", newline="") as f:
        f.write("\r\n".join(cfg) + "\r\n")
    degs = np.linspace(0, 360, na, endpoint=False) + 0.125
    if use_ang:
        with open(os.path.join(tmp, "scan.ang"), "w") as f:
            f.write("Projection: Angle\n")
            for i, d in enumerate(degs):
                f.write(f"{i + 1}: {d:.4f}\n")
    else:
        with open(os.path.join(tmp, "_ctdata.txt"), "w") as f:
            f.write("Proj Angle Time\n0 0 0\nProj Angle(deg) Time\n")
            for i, d in enumerate(degs):
                f.write(f"{i + 1}\t{d:.4f}\t{0.1 * i:.1f}\r\n")
    for i in range(na):
        _pil16(os.path.join(tmp, f"scan{sep}{i + 1:04d}.tif"), imgs[i], big)
    return imgs, degs


@pytest.mark.parametrize("use_ang,big", [(False, False), (True, True)])
def test_load_xtek_layout_angles_and_log(tmp_path, use_ang, big):
    na, nh, nvd = 5, 7, 6
    imgs, degs = _write_scan(str(tmp_path), na, nh, nvd, use_ang=use_ang, big=big)
    pix, ang, sx, dx, ph, pv, mr = F.load_xtek(str(tmp_path / "scan.xtekct"))
    # cone v alignment: 6 rows -> 2 blocks of 4 -> 8 rows, data at offset 1
    assert pix.shape == (na, nh, 8) and pix.dtype == np.float32
    assert (sx, dx, ph, pv, mr) == (120.5, 700.25, 0.2, 0.2, 11.75)
    want_deg = 10.0 + np.round(degs, 4)
    if use_ang:
        want_deg = want_deg[::-1]
    assert np.allclose(ang, want_deg, atol=1e-4)
    assert np.all(pix[:, :, 0] == 0) and np.all(pix[:, :, 7] == 0)
    # pixel (a, h, v) is image row nvd - 1 - v, column h
    raw = imgs[:, ::-1, :].transpose(0, 2, 1).astype(np.float64)
    wl, sc = 4000.0, 2.5
    top = raw.max() if raw.max() > wl else wl
    top -= wl * sc / 100.0
    d = (raw - wl * sc / 100.0).astype(np.float32).astype(np.float64)
    with np.errstate(invalid="ignore"):
        want = np.where(d < 1.0, -math.log(1e-5 / top), -np.log(d / top)).astype(np.float32)
    assert np.array_equal(pix[:, :, 1:7], want)


def test_read_xtekct_geometry_and_errors(tmp_path):
    _write_scan(str(tmp_path), 3, 4, 5)
    c = F.read_xtekct(str(tmp_path / "scan.xtekct"))
    assert c["source_x"] == -120.5 and c["detector_x"] == 700.25 - 120.5
    assert c["sep"] == "_" and c["name"] == "scan"
    assert np.allclose(c["h_pixels"], (np.arange(4) - 1.5) * 0.4)
    assert c["nv_aligned"] == 8 and np.allclose(c["v_pixels"], (np.arange(8) - 3.5) * 0.5)
    text = open(tmp_path / "scan.xtekct").read().replace("WhiteLevel", "Whitelevel")
    (tmp_path / "bad.xtekct").write_text(text)
    with pytest.raises(F.FormatError, match="all geometry"):
        F.read_xtekct(str(tmp_path / "bad.xtekct"))
    (tmp_path / "hdr.xtekct").write_text("[XTek]\n")
    with pytest.raises(F.FormatError, match="header"):
        F.read_xtekct(str(tmp_path / "hdr.xtekct"))


def test_write_results_tiff_clamped_and_stretched(tmp_path):
    v = np.array([-0.5, 0.0, 0.5, 0.999, 1.0, 2.0], np.float32).reshape(2, 3, 1)
    base = str(tmp_path / "r")
    F.write_results(base, v, (0, 0, 0), (1, 1, 1), fmt=F.UNSIGNED_BYTE_TIFF, clamp=True)
    img = np.array(Image.open(base + "_0001.tif"))[::-1]
    # image row j, column i holds voxel (i, j)
    assert np.array_equal(img.T.ravel(), [0, 0, 127, 254, 255, 255])
    F.write_results(base, v, (0, 0, 0), (1, 1, 1), offset=4, fmt=F.UNSIGNED_SHORT_TIFF)
    img = np.array(Image.open(base + "_0005.tif"))[::-1].T.ravel()
    assert img[0] == 0 and img[-1] == 65535
    assert img[2] == int(np.float32(1.0) * np.float32(65535 / np.float32(2.5)))


def test_write_results_bgs_and_native_dump(tmp_path):
    v = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    base = str(tmp_path / "b")
    F.write_results(base, v, (-1.0, -1.5, -2.0), (1.0, 1.0, 1.0), fmt=F.BGS_FLOAT_DUMP)
    raw = open(base + ".dat", "rb").read()
    hdr = b"2 3 4\n   -0.500000    -1.000000    -1.500000\n    1.000000 0.0 0.0\n" \
          b"0.0     1.000000 0.0\n0.0 0.0     1.000000\nImage\n"
    assert raw.startswith(hdr)
    data = np.frombuffer(raw[len(hdr):], "<f4").reshape(4, 3, 2)
    assert np.array_equal(data, v.transpose(2, 1, 0))
    base2 = str(tmp_path / "n")
    F.write_results(base2, v[:, :, :2], None, None, fmt=F.NATIVE_DUMP)
    F.write_results(base2, v[:, :, 2:], None, None, offset=2, fmt=F.NATIVE_DUMP)
    out = np.frombuffer(open(base2 + ".dat", "rb").read(), "<f4")
    assert np.array_equal(out, np.concatenate([v[:, :, :2].ravel(), v[:, :, 2:].ravel()]))
