"""C++ callers: the reference-shaped facade (include/ccpi_hip.hpp) drives
FP, BP, CGLS and cgls_tikhonov through the C ABI (GPU test; binary built by
build()).  The binary writes every result array; each is compared element
by element with the oracle on the same inputs."""

import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
N, NA, NH, NV = 10, 12, 16, 14


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_cxx_facade_matches_oracle(tmp_path):
    exe = os.path.join(HERE, "cxx", "facade_check")
    assert os.path.exists(exe), "build() compiles tests/cxx/facade_check"
    out = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.split()[:2] == ["facade", "ok"]

    def load(name, dtype, shape):
        return np.fromfile(tmp_path / name, dtype=dtype).reshape(shape)

    vol, sino = (N, N, N), (NA, NH, NV)
    og = oracle.Geometry(2 * np.pi * np.arange(NA) / NA, (np.arange(NH) - 7.5) * 1.5,
                         (np.arange(NV) - 6.5) * 1.5, (-5, -5, -5), (1, 1, 1), vol,
                         cone=True, source=(-30, 0, 0), det_x=30.0)
    # the binary's input volume: x[q] = (q * 37 mod 101) / 100
    x = ((np.arange(N ** 3, dtype=np.int64) * 37 % 101).astype(np.float32)
         / np.float32(100.0)).reshape(vol)
    y = load("fp.f32", np.float32, sino)
    assert _rel(y, og.fp(x)) < 1e-6
    b = load("bp.f32", np.float32, vol)
    assert _rel(b, og.bp(y)) < 1e-5
    xc = load("cgls_x.f32", np.float32, vol)
    xo, _, no = oracle.cgls(og, y, 3)
    assert _rel(xc, xo) < 1e-4
    assert np.allclose(load("cgls_norms.f64", np.float64, (3,)),
                       oracle.cgls(og, y, 3, dtype=np.float64)[2], rtol=2e-4)
    xt = load("tik_x.f32", np.float32, vol)
    xto, _, _ = oracle.cgls_reg(og, y, 3, oracle.TIKHONOV, 0.2)
    assert _rel(xt, xto) < 1e-4
    assert np.allclose(load("tik_norms.f64", np.float64, (3,)),
                       oracle.cgls_reg(og, y, 3, oracle.TIKHONOV, 0.2, dtype=np.float64)[2],
                       rtol=2e-4)
