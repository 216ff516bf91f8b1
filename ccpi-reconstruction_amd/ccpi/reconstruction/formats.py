"""Input and output formats around the reconstruction path.

  * Nikon XTek scans: the ``.xtekct`` configuration, the ``_ctdata.txt`` /
    ``.ang`` angle lists and the 16-bit projection TIFFs
    (Core/src/Readers/xtek.cpp:175-366, 460-517; Readers/tiff.cpp:22-116),
    with the reference's white-level normalisation and -log;
  * result writers: 8/16-bit TIFF slices, the raw float dump and the
    ``.bgs`` float dump (Core/src/results.cpp:23-228, Readers/tiff.cpp:
    118-176).

TIFF is read and written here without libtiff: baseline strips, either
byte order; the reader decodes uncompressed, LZW, Deflate and PackBits
strips (with the horizontal predictor), the writer writes uncompressed
strips.  Host-side numpy throughout; the arrays
feed the device path through the ``conebeam`` / ``parallelbeam`` calls.
"""

from __future__ import annotations

import math
import os
import re
import struct
import zlib

import numpy as np

from .instruments import calc_v_alignment

# ---------------------------------------------------------------------------
# TIFF

_TYPE_SIZE = {1: 1, 2: 1, 3: 2, 4: 4, 5: 8, 6: 1, 7: 1, 8: 2, 9: 4, 10: 8, 11: 4, 12: 8,
              16: 8}
_TYPE_FMT = {1: "B", 3: "H", 4: "I", 6: "b", 7: "B", 8: "h", 9: "i", 16: "Q"}


class FormatError(ValueError):
    pass


def _tiff_tags(buf: bytes):
    if buf[:2] == b"II":
        bo = "<"
    elif buf[:2] == b"MM":
        bo = ">"
    else:
        raise FormatError("not a TIFF file")
    if struct.unpack(bo + "H", buf[2:4])[0] != 42:
        raise FormatError("not a classic TIFF file")
    ifd = struct.unpack(bo + "I", buf[4:8])[0]
    (n,) = struct.unpack(bo + "H", buf[ifd:ifd + 2])
    tags = {}
    for e in range(n):
        off = ifd + 2 + 12 * e
        tag, typ, cnt = struct.unpack(bo + "HHI", buf[off:off + 8])
        if typ not in _TYPE_FMT:
            continue
        size = _TYPE_SIZE[typ] * cnt
        if size <= 4:
            raw = buf[off + 8:off + 8 + size]
        else:
            (ptr,) = struct.unpack(bo + "I", buf[off + 8:off + 12])
            raw = buf[ptr:ptr + size]
        tags[tag] = struct.unpack(bo + _TYPE_FMT[typ] * cnt, raw)
    return bo, tags


def _unpackbits(raw: bytes) -> bytes:
    """PackBits (TIFF 6.0 section 9): a header byte n, then n + 1 literal
    bytes (0 <= n <= 127), or one byte repeated 1 - n times (-127 <= n <= -1);
    n = -128 is a no-op."""
    out = bytearray()
    i, n = 0, len(raw)
    while i < n:
        h = raw[i]
        i += 1
        if h < 128:
            out += raw[i:i + h + 1]
            i += h + 1
        elif h > 128:
            if i < n:
                out += bytes([raw[i]]) * (257 - h)
            i += 1
    return bytes(out)


def _unlzw(raw: bytes) -> bytes:
    """TIFF LZW (TIFF 6.0 section 13): MSB-first codes of 9 to 12 bits,
    256 clears the table, 257 ends the strip; the code width grows one code
    early (at table sizes 511, 1023, 2047)."""
    table = [bytes([q]) for q in range(256)] + [b"", b""]
    out = bytearray()
    bits = 0
    nbits = 0
    width = 9
    prev = None
    pos = 0
    n = len(raw)
    while True:
        while nbits < width and pos < n:
            bits = (bits << 8) | raw[pos]
            pos += 1
            nbits += 8
        if nbits < width:
            break
        code = (bits >> (nbits - width)) & ((1 << width) - 1)
        nbits -= width
        bits &= (1 << nbits) - 1
        if code == 257:
            break
        if code == 256:
            table = table[:258]
            width = 9
            prev = None
            continue
        if prev is None:
            entry = table[code]
        elif code < len(table):
            entry = table[code]
            table.append(prev + entry[:1])
        elif code == len(table):
            entry = prev + prev[:1]
            table.append(entry)
        else:
            raise FormatError("corrupt LZW strip")
        out += entry
        prev = entry
        if len(table) + 1 >= (1 << width) and width < 12:
            width += 1
    return bytes(out)


def _decode_strip(raw: bytes, compression: int) -> bytes:
    if compression == 1:
        return raw
    if compression == 5:
        return _unlzw(raw)
    if compression in (8, 32946):
        return zlib.decompress(raw)
    if compression == 32773:
        return _unpackbits(raw)
    raise FormatError("TIFF compression %d is not supported" % compression)


def read_tiff16(path: str) -> np.ndarray:
    """One 16-bit grey-level TIFF as uint16 [rows][width], with the checks
    of read_tiff (tiff.cpp:22-116): strips (not tiles), 16 bits per sample,
    min-is-black, MSB-to-LSB fill order.  Strips may be uncompressed, LZW,
    Deflate or PackBits, with or without the horizontal predictor (the
    reference decodes them through libtiff)."""
    with open(path, "rb") as f:
        buf = f.read()
    bo, t = _tiff_tags(buf)
    if 322 in t or 324 in t:
        raise FormatError("tiled image not supported in XTek reader")
    if t.get(258, (1,))[0] != 16:
        raise FormatError("TIFF is not 16bit data")
    if 262 not in t:
        raise FormatError("Error getting TIFF photometric info")
    if t[262][0] != 1:
        raise FormatError("TIFF photometric type not supported by XTek reader")
    if t.get(266, (1,))[0] != 1:
        raise FormatError("TIFF is not MSB to LSB")
    if t.get(277, (1,))[0] != 1:
        raise FormatError("TIFF has more than one sample per pixel")
    compression = t.get(259, (1,))[0]
    predictor = t.get(317, (1,))[0]
    if predictor not in (1, 2):
        raise FormatError("TIFF predictor %d is not supported" % predictor)
    width, rows = t[256][0], t[257][0]
    data = b"".join(_decode_strip(buf[o:o + c], compression) for o, c in zip(t[273], t[279]))
    if len(data) % 2:
        raise FormatError("Odd number of bytes for 16 bit image")
    img = np.frombuffer(data, dtype=np.dtype(bo + "u2"))
    if img.size != width * rows:
        raise FormatError("Image size mismatch")
    img = img.reshape(rows, width).astype(np.uint16)
    if predictor == 2:
        # horizontal differencing: each sample is stored minus its left
        # neighbour, modulo 2^16
        img = np.cumsum(img, axis=1, dtype=np.uint16)
    return img


def write_tiff(path: str, data: np.ndarray) -> None:
    """write_tiff (tiff.cpp:118-176): [ny][nx] uint8 or uint16, one row per
    strip, rows stored bottom-up (file row r holds data row ny - 1 - r),
    min-is-black, little-endian."""
    a = np.asarray(data)
    if a.dtype not in (np.uint8, np.uint16) or a.ndim != 2:
        raise FormatError("write_tiff takes a 2-D uint8 or uint16 array")
    ny, nx = a.shape
    rows = np.ascontiguousarray(a[::-1]).astype(a.dtype.newbyteorder("<"))
    pix = rows.tobytes()
    row_bytes = nx * a.itemsize
    ntags = 10
    ifd_off = 8
    ifd_size = 2 + 12 * ntags + 4
    offs_off = ifd_off + ifd_size
    cnts_off = offs_off + 4 * ny
    data_off = cnts_off + 4 * ny
    entries = [
        (256, 4, 1, nx), (257, 4, 1, ny), (258, 3, 1, 8 * a.itemsize), (259, 3, 1, 1),
        (262, 3, 1, 1), (273, 4, ny, offs_off if ny > 1 else data_off), (277, 3, 1, 1),
        (278, 4, 1, 1), (279, 4, ny, cnts_off if ny > 1 else row_bytes), (284, 3, 1, 1),
    ]
    out = bytearray(b"II" + struct.pack("<HI", 42, ifd_off))
    out += struct.pack("<H", ntags)
    for tag, typ, cnt, val in entries:
        if typ == 3 and cnt == 1:
            out += struct.pack("<HHIHH", tag, typ, cnt, val, 0)
        else:
            out += struct.pack("<HHII", tag, typ, cnt, val)
    out += struct.pack("<I", 0)
    out += struct.pack(f"<{ny}I", *[data_off + r * row_bytes for r in range(ny)])
    out += struct.pack(f"<{ny}I", *([row_bytes] * ny))
    out += pix
    with open(path, "wb") as f:
        f.write(out)


# ---------------------------------------------------------------------------
# XTek scans

_NUM = re.compile(r"\s*[-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?")


def _atof(s: str) -> float:
    m = _NUM.match(s)
    return float(m.group(0)) if m else 0.0


def _atoi(s: str) -> int:
    m = re.match(r"\s*[-+]?\d+", s)
    return int(m.group(0)) if m else 0


def read_xtekct(filename: str) -> dict:
    """Parse an .xtekct file as read_config_file does (xtek.cpp:175-299):
    fixed key prefixes, values read C-style from after the '=', eleven
    required entries.  Returns the fields plus the geometry derived from
    them (source x, detector x, h / v pixel positions, angles in radians)."""
    path, name = os.path.split(filename)
    with open(filename, "r", errors="replace", newline="") as f:
        lines = f.read().splitlines()
    if not lines or not lines[0].startswith("[XTekCT]"):
        raise FormatError("Incorrect header on xtekct file")
    c = dict(src_to_object=0.0, src_to_detector=0.0, mask_radius=0.0, n_h=0, n_v=0,
             pixel_h=0.0, pixel_v=0.0, n_angles=0, initial_angle=0.0, white_level=0.0,
             scattering=0.0, coef=(1.0, 0.0, 0.0, 0.0, 0.0), scale=1.0, name="", sep="")
    coef = {"CoefX4": 0.0, "CoefX3": 0.0, "CoefX2": 1.0, "CoefX1": 0.0, "CoefX0": 0.0}
    nd = 0
    keys = [("SrcToObject", "src_to_object", _atof), ("SrcToDetector", "src_to_detector", _atof),
            ("MaskRadius", "mask_radius", _atof), ("DetectorPixelsX", "n_h", _atoi),
            ("DetectorPixelsY", "n_v", _atoi), ("DetectorPixelSizeX", "pixel_h", _atof),
            ("DetectorPixelSizeY", "pixel_v", _atof), ("Projections", "n_angles", _atoi),
            ("InitialAngle", "initial_angle", _atof), ("WhiteLevel", "white_level", _atof)]
    for line in lines[1:]:
        line = line[:255]
        for key, field, conv in keys:
            if line.startswith(key):
                c[field] = conv(line[len(key) + 1:])
                nd += 1
                break
        else:
            if line.startswith("InputSeparator"):
                ch = line[15] if len(line) > 15 else "\0"
                if ch != "\r" and ch != "\0":
                    c["sep"] = ch
            elif line.startswith("Scattering"):
                c["scattering"] = _atof(line[11:])
            elif line[:6] in coef:
                coef[line[:6]] = _atof(line[7:])
            elif line.startswith("Scale"):
                c["scale"] = _atof(line[6:])
            elif line.startswith("Name"):
                c["name"] = line[5:].rstrip()
                nd += 1
    c["coef"] = tuple(coef[k] for k in ("CoefX4", "CoefX3", "CoefX2", "CoefX1", "CoefX0"))
    if nd != 11:
        raise FormatError("Failed to locate all geometry data in xtekct file")
    if c["name"] == "":
        raise FormatError("No name found for tiff files")
    if not (c["n_h"] > 0 and c["n_v"] > 0 and c["n_angles"] > 0):
        raise FormatError("Negative pixel values")
    nv, voff = calc_v_alignment(c["n_v"], 1, True)
    nh = c["n_h"]
    h = -((nh - 1) * c["pixel_h"] / 2.0) + np.arange(nh) * c["pixel_h"]
    v = -((nv - 1) * c["pixel_v"] / 2.0) + np.arange(nv) * c["pixel_v"]
    c.update(path=path, source_x=-c["src_to_object"],
             detector_x=c["src_to_detector"] - c["src_to_object"], h_pixels=h, v_pixels=v,
             nv_aligned=nv, v_offset=voff,
             angles=read_angles(path, c["name"], c["initial_angle"], c["n_angles"]))
    return c


def read_angles(path: str, basename: str, init_angle: float, n: int) -> np.ndarray:
    """read_angles (xtek.cpp:301-352): the second column of _ctdata.txt
    after a 9-token header, else the '<index>: <angle>' lines of
    <basename>.ang (first line skipped) in reverse order; degrees plus the
    initial angle, returned in radians."""
    ct = os.path.join(path, "_ctdata.txt")
    out = np.zeros(n, np.float64)
    if os.path.exists(ct):
        tok = open(ct, "r", errors="replace").read().split()
        vals = [_atof(tok[9 + 3 * i + 1]) for i in range(n)]
        for i in range(n):
            out[i] = math.pi * (init_angle + vals[i]) / 180.0
        return out
    ang = os.path.join(path, basename + ".ang")
    if not os.path.exists(ang):
        raise FormatError("Error opening ctdata file")
    lines = open(ang, "r", errors="replace").read().splitlines()[1:]
    pat = re.compile(r"\s*[-+0-9.eE]+\s*\S\s*([-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?)")
    for i in range(n):
        m = pat.match(lines[i])
        tmp = float(m.group(1)) if m else 0.0
        out[n - i - 1] = math.pi * (init_angle + tmp) / 180.0
    return out


def read_xtek_images(cfg: dict) -> np.ndarray:
    """read_images (xtek.cpp:460-517): <name><sep>NNNN.tif per angle into
    [angle][h][v] with v flipped (row 0 of a TIFF is the top) and centred in
    the aligned rows; then the white-level / scattering normalisation and
    -log, rows outside the data set to 0."""
    na, nh, nv = cfg["n_angles"], cfg["n_h"], cfg["nv_aligned"]
    nvd, voff = cfg["n_v"], cfg["v_offset"]
    base = os.path.join(cfg["path"], cfg["name"] + cfg["sep"])
    pix = np.zeros((na, nh, nv), np.float32)
    for i in range(na):
        img = read_tiff16(f"{base}{i + 1:04d}.tif")
        if img.shape[1] != nh:
            raise FormatError("Image width mismatch")
        if img.size != nh * nvd:
            raise FormatError("Image size mismatch")
        pix[i, :, voff:voff + nvd] = img.reshape(nvd, nh)[::-1].T
    data = pix[:, :, voff:voff + nvd]
    wl, sc = cfg["white_level"], cfg["scattering"]
    max_v = float(data.max()) if data.size else 0.0
    if not max_v > wl:
        max_v = wl
    max_v -= wl * sc / 100.0
    d = (data.astype(np.float64) - wl * sc / 100.0).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        low = np.float32(-math.log(0.00001 / max_v)) if max_v > 0 else np.float32(np.nan)
        out = np.where(d < 1.0, low, (-np.log(d.astype(np.float64) / max_v)).astype(np.float32))
    pix[:, :, voff:voff + nvd] = out
    return pix


def load_xtek(filename: str):
    """conebeam_load_xtek (conebeam_wrapper.cpp:313-344): (pixels
    [angle][h][v] after -log, angles in degrees (float32),
    source_to_sample, source_to_detector, 0.2, 0.2, mask_radius).  The
    pixel sizes returned are the reference's fixed 0.2 mm, not the file's;
    the centre search that ends read_images only adjusts the reader's own
    geometry, which is not returned, so it is not run here."""
    cfg = read_xtekct(filename)
    pix = read_xtek_images(cfg)
    ang = (cfg["angles"] * 180.0 / math.pi).astype(np.float32)
    sx = -cfg["source_x"]
    dx = cfg["detector_x"] + sx
    return pix, ang, sx, dx, 0.2, 0.2, cfg["mask_radius"]


# ---------------------------------------------------------------------------
# results

NO_OUTPUT, UNSIGNED_BYTE_TIFF, UNSIGNED_SHORT_TIFF, NATIVE_DUMP, SIGNED_SHORT_TIFF, \
    BGS_FLOAT_DUMP = range(6)


def write_results(basename: str, voxels: np.ndarray, voxel_origin, voxel_size, offset: int = 0,
                  nz_voxels: int | None = None, fmt: int = UNSIGNED_SHORT_TIFF,
                  clamp: bool = False) -> None:
    """write_results (results.cpp:23-56) for a [nx][ny][nz] float volume."""
    v = np.asarray(voxels, np.float32)
    if fmt == NO_OUTPUT:
        return
    if fmt in (UNSIGNED_BYTE_TIFF, UNSIGNED_SHORT_TIFF, SIGNED_SHORT_TIFF):
        top, width = {UNSIGNED_BYTE_TIFF: (255, 8), UNSIGNED_SHORT_TIFF: (65535, 16),
                      SIGNED_SHORT_TIFF: (32767, 16)}[fmt]
        _write_as_tiff(basename, v, offset, top, width, clamp)
    elif fmt == NATIVE_DUMP:
        _write_real(basename, v, offset)
    elif fmt == BGS_FLOAT_DUMP:
        _write_bgs(basename, v, voxel_origin, voxel_size, offset,
                   v.shape[2] if nz_voxels is None else nz_voxels)
    else:
        raise FormatError("Unknown output format")


def _write_as_tiff(basename, v, offset, top, width, clamp):
    """write_as_tiff (results.cpp:58-133): slice k -> <base>_NNNN.tif as
    [ny][nx]; clamp maps [0, 1) to [0, top) by truncation, otherwise the
    volume's range is stretched to [0, top] (float scale)."""
    dt = np.uint8 if width == 8 else np.uint16
    ftop = np.float32(top)
    if not clamp:
        vmin, vmax = np.float32(v.min()), np.float32(v.max())
        scale = np.float32(ftop / np.float32(vmax - vmin))
    for k in range(v.shape[2]):
        s = v[:, :, k].T
        if clamp:
            q = np.where(s < 0.0, np.float32(0), np.where(s >= 1.0, ftop, s * ftop))
        else:
            q = (s - vmin) * scale
        img = np.trunc(q).astype(np.int64).astype(dt)
        write_tiff(f"{basename}_{offset + k + 1:04d}.tif", img)


def _write_real(basename, v, offset):
    """write_real (results.cpp:136-158): raw float [nx][ny][nz] appended to
    <base>.dat.  The reference seeks to the block's offset first, but the
    file is open in append mode, where every write goes to the end; blocks
    written in order land at their offsets either way."""
    with open(basename + ".dat", "ab") as f:
        f.write(np.ascontiguousarray(v).astype("<f4").tobytes())


def _write_bgs(basename, v, origin, size, offset, nz_voxels):
    """write_bgs (results.cpp:160-228): a text header (dimensions, the
    centre of the first voxel, the axes, 'Image') when offset is 0, then
    float data ordered [z][y][x], appended block by block."""
    name = basename + ".dat"
    with open(name, "wb" if offset == 0 else "ab") as f:
        if offset == 0:
            sh = [origin[q] + size[q] / 2.0 for q in range(3)]
            hdr = (f"{v.shape[0]} {v.shape[1]} {nz_voxels}\n"
                   f"{sh[0]:12.6f} {sh[1]:12.6f} {sh[2]:12.6f}\n"
                   f"{size[0]:12.6f} 0.0 0.0\n0.0 {size[1]:12.6f} 0.0\n"
                   f"0.0 0.0 {size[2]:12.6f}\nImage\n")
            f.write(hdr.encode())
        f.write(np.ascontiguousarray(v.transpose(2, 1, 0)).astype("<f4").tobytes())
