"""Cone geometries that select each backprojection kernel instance, and the
host's choice of the BP accumulation layout restated (shared by the CPU
layout test and the GPU parity test of every BP path)."""

import math

import numpy as np

from geometries import CB, centred

A12 = np.linspace(0, 2 * np.pi, 12, endpoint=False) + 0.1

CASES = {
    "s1": (CB((10, 10, 8), (1, 1, 1), np.linspace(0, 2 * np.pi, 12, endpoint=False) + 0.1,
              centred(20, 1.6), centred(8, 3.0), 40.0, 80.0), (1, False, 1)),
    # three 128-row blocks: one under the source row, one across it, one
    # above it with padded rows; floor-form and main rows share the first
    "s2_blocks": (CB((10, 10, 120), (1, 1, 1), A12, centred(20, 1.6), centred(300, 1.6),
                     40.0, 80.0), (2, False, 1)),
    "s3": (CB((10, 10, 6), (1, 1, 1), np.linspace(0, 2 * np.pi, 12, endpoint=False) + 0.1,
              centred(20, 1.6), centred(20, 0.95), 40.0, 80.0), (3, False, 1)),
    "s4": (CB((10, 10, 6), (1, 1, 1), np.linspace(0, 2 * np.pi, 12, endpoint=False) + 0.1,
              centred(20, 1.6), centred(24, 0.75), 40.0, 80.0), (4, False, 1)),
    "phased_s8_small": (CB((8, 8, 4), (1, 1, 1), np.linspace(0, 2 * np.pi, 10, endpoint=False),
                           centred(16, 1.6), centred(40, 0.4), 40.0, 80.0), (8, True, 1)),
    "phased_s8_offz": (CB((8, 8, 4), (1, 1, 1), np.linspace(0, 2 * np.pi, 10, endpoint=False),
                          centred(16, 1.6), centred(40, 0.4), 40.0, 80.0, source_z=1.7),
                       (8, True, 1)),
    "phased_s8_main": (CB((12, 12, 200), (1, 1, 1), A12, centred(20, 1.6), centred(460, 0.5),
                          20.0, 40.0), (8, True, 1)),
    "phased_s8_blocks": (CB((12, 12, 160), (1, 1, 0.25), A12, centred(20, 1.6),
                            centred(900, 0.17), 20.0, 40.0), (8, True, 1)),
    "phased_s12": (CB((12, 12, 40), (1, 1, 1), A12, centred(20, 1.6), centred(700, 0.25),
                      32.0, 64.0), (12, True, 1)),
    "phased_s12_2ph": (CB((12, 12, 40), (1, 1, 1), A12, centred(20, 1.6), centred(700, 0.25),
                          24.0, 48.0), (12, True, 2)),
    "phased_s16": (CB((12, 12, 40), (1, 1, 1), A12, centred(20, 1.6), centred(1000, 0.25),
                      24.0, 48.0), (16, True, 1)),
    "phased_s8_2ph": (CB((12, 12, 40), (1, 1, 1), A12, centred(20, 1.6), centred(300, 0.25),
                         24.0, 48.0), (8, True, 2)),
    "phased_s16_2ph": (CB((12, 12, 40), (1, 1, 1), A12, centred(20, 1.6), centred(1000, 0.2),
                          24.0, 48.0), (16, True, 2)),
    "phased_s4_lds": (CB((4, 4, 1500), (1, 1, 0.05), np.linspace(0, 2 * np.pi, 8, endpoint=False),
                         centred(10, 1.2), centred(3000, 0.05), 40.0, 80.0), (4, True, 1)),
    # source 0.002 mm outside the volume's corner circle: up to ~2800 rows
    # per slice near the corners, 64 lane phases (one lane per round); a
    # bound clamped at walk parameter 1e-3 would give 33 phases of 16 slots,
    # 528 rows apart, on 640 rows that all fall within two slices there
    "phased_corner": (CB((8, 8, 6), (1, 1, 0.5), np.deg2rad([0, 45, 100, 135, 225, 260, 315]),
                         centred(14, 1.0), centred(640, 1.0), math.sqrt(32) + 0.002,
                         2 * (math.sqrt(32) + 0.002)), (12, True, 64)),
}


# With the source inside the volume's top slice the reference's b2D is not
# the transpose of its f2D (the oracle's own <Ax, y> and <x, A^T y> differ
# by 31 % for this case; test_oracle_pins maps the heights): parity only.
NOT_ADJOINT = {"phased_s8_offz"}


def bp_path(g):
    """(row stride, phased, phases) as the host chooses them (cone.hip plan
    build)."""
    nx, ny, nz = g.shape
    vs, o = g.vsize, g.origin
    R = max(math.hypot(o[0] + cx * nx * vs[0], o[1] + cy * ny * vs[1])
            for cx in (0, 1) for cy in (0, 1))
    Lmax = max(math.hypot(g.det_x - g.source[0], g.h_pix[i] - g.source[1]) for i in (0, -1))
    amin = (math.hypot(g.source[0], g.source[1]) - R) / Lmax
    want = max(1, math.ceil(min(vs[2] / (amin * np.diff(g.v_pix).min()) * 1.05, 1e6)))
    s = min(16, want)
    pad = (nz + 8 + 3) & ~3  # four trash cells on each side
    lds = 4 * (pad + 4 * s * pad)
    phased = not (want <= 16 and s <= 4 and lds <= 64 * 1024)
    if not phased:
        return s, False, 1
    nv = len(g.v_pix)

    def phases(s):
        return min(64, -(-want // s))

    def cost(s):
        return -(-nv // (64 * s)) * 64 * s * (1.0 + 0.11 * (phases(s) - 1))

    s = 4
    if want > 4:
        s = 8
        for c in (12, 16):          # the smaller S on a tie
            if cost(c) < cost(s):
                s = c
    return s, True, phases(s)


