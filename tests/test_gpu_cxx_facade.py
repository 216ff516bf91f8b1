"""C++ callers: the reference-shaped facade (include/ccpi_hip.hpp) drives
FP, BP, CGLS and cgls_tikhonov through the C ABI (GPU test; binary built by
build())."""

import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_cxx_facade_matches_oracle():
    exe = os.path.join(HERE, "cxx", "facade_check")
    assert os.path.exists(exe), "build() compiles tests/cxx/facade_check"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    parts = out.stdout.split()
    assert parts[:2] == ["facade", "ok"]
    fps, bps, norm0, tik0 = (float(q) for q in parts[2:6])
    n, na, nh, nv = 10, 12, 16, 14
    og = oracle.Geometry(2 * np.pi * np.arange(na) / na, (np.arange(nh) - 7.5) * 1.5,
                         (np.arange(nv) - 6.5) * 1.5, (-5, -5, -5), (1, 1, 1), (n, n, n),
                         cone=True, source=(-30, 0, 0), det_x=30.0)
    y = og.fp(np.ones((n, n, n), np.float64))
    assert abs(fps - y.sum()) < 1e-5 * y.sum()
    x = og.bp(y.astype(np.float32).astype(np.float64))
    assert abs(bps - x.sum()) < 1e-5 * x.sum()
    _, _, norms = oracle.cgls(og, y.astype(np.float32), 3, dtype=np.float64)
    assert abs(norm0 - norms[0]) < 1e-4 * norms[0]
    _, _, tn = oracle.cgls_reg(og, y.astype(np.float32), 3, oracle.TIKHONOV, 0.2,
                               dtype=np.float64)
    assert abs(tik0 - tn[0]) < 1e-4 * tn[0]
