"""The cone BP's accumulation layout as the library chooses it (CPU test:
ccpi_cone_bp_layout is a host computation and touches no device).

The BP writes each detector row's contribution into an LDS copy of the
voxel column with plain read-add-writes, so the lanes of one write round
must fall on distinct slices.  That holds when the rows of one round are
at least as many rows apart as the most detector rows that share a slice
anywhere in the volume.  Here that maximum is found by brute force: for
every (angle, h) ray the walk parameter where it enters the volume's xy
square (slab method), where adjacent rows are alpha * pitch apart in z.
"""

import ctypes as C
import math

import numpy as np
import pytest

from bp_cases import CASES, bp_path
from ccpi.reconstruction import _lib
from ccpi.reconstruction.projectors import Geometry


def _geom(g):
    return Geometry(g.angles, g.h_pix, g.v_pix, g.origin, g.vsize, g.shape, cone=True,
                    source=g.source, detector_x=g.det_x)


def _layout(g):
    L = _lib.lib()
    out = [C.c_int() for _ in range(4)]
    _lib.check(L.ccpi_cone_bp_layout(C.byref(_geom(g).c_struct()), *[C.byref(q) for q in out]))
    return tuple(q.value for q in out)   # stride, phased, phases, rows bound


def _max_rows_per_slice(g, angles):
    """Largest vsize_z / (alpha_entry * pitch) over the rays that cross the
    volume's xy square."""
    nx, ny, _ = g.shape
    x0, y0 = g.origin[0], g.origin[1]
    xn, yn = x0 + nx * g.vsize[0], y0 + ny * g.vsize[1]
    c, s = np.cos(angles)[:, None], np.sin(angles)[:, None]
    sx, sy = g.source[0], g.source[1]
    p1x, p1y = c * sx - s * sy, s * sx + c * sy
    h = g.h_pix[None, :]
    p2x, p2y = c * g.det_x - s * h, s * g.det_x + c * h
    dx, dy = p2x - p1x, p2y - p1y
    with np.errstate(divide="ignore", invalid="ignore"):
        ax0, axn = (x0 - p1x) / dx, (xn - p1x) / dx
        ay0, ayn = (y0 - p1y) / dy, (yn - p1y) / dy
        lo = np.maximum(np.maximum(np.minimum(ax0, axn), np.minimum(ay0, ayn)), 0.0)
        hi = np.minimum(np.minimum(np.maximum(ax0, axn), np.maximum(ay0, ayn)), 1.0)
    hit = np.isfinite(lo) & np.isfinite(hi) & (hi > lo)
    pitch = np.diff(g.v_pix).min()
    return float((g.vsize[2] / (lo[hit] * pitch)).max())


@pytest.mark.parametrize("name", list(CASES))
def test_layout_matches_restated_choice(name):
    g, want_path = CASES[name]
    stride, phased, phases, _ = _layout(g)
    assert (stride, bool(phased), phases) == bp_path(g) == want_path


@pytest.mark.parametrize("name", list(CASES))
def test_rounds_fall_on_distinct_slices(name):
    g, _ = CASES[name]
    stride, phased, phases, bound = _layout(g)
    dense = np.linspace(0, 2 * np.pi, 3600, endpoint=False)
    worst = max(_max_rows_per_slice(g, g.angles), _max_rows_per_slice(g, dense))
    assert bound >= worst
    if phased:
        assert phases * stride >= bound or phases == 64
    else:
        assert stride >= bound


def test_source_at_corner_circle():
    """A source 0.002 mm outside the corner circle: the bound reaches the
    thousands of rows that share a slice near the corner facing the source,
    so the phased form runs one lane per round."""
    g, _ = CASES["phased_corner"]
    stride, phased, phases, bound = _layout(g)
    worst = _max_rows_per_slice(g, g.angles)
    assert worst > 2000
    assert bound >= worst and phased == 1 and phases == 64


def test_layout_rejects_parallel_geometry():
    from geometries import pb_small
    g = pb_small()
    G = Geometry(g.angles, g.h_pix, g.v_pix, g.origin, g.vsize, g.shape)
    out = [C.c_int() for _ in range(4)]
    st = _lib.lib().ccpi_cone_bp_layout(C.byref(G.c_struct()), *[C.byref(q) for q in out])
    assert st == -1
