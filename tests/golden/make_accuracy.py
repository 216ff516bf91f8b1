"""Writes the accuracy fixtures of the cone-beam CGLS workloads (run in a
container with many host cores; minutes to an hour per fixture):

  python tests/golden/make_accuracy.py cone512      # 512^3, 10 iterations
  python tests/golden/make_accuracy.py cone256x100  # 768^3 study at 256^3, 100 iterations

For each workload: the phantom and the Poisson-noisy sinogram are built on
the host (workloads.ellipsoid_phantom_host / poisson_log_host, numpy PCG64
seeds), the noise-free sinogram is the float oracle's cone f2D of the
phantom, and CGLS runs from x = 0 in the oracle's float instantiation (the
reference's arithmetic: float weights, float partial sums and alpha) and in
its FP64 instantiation.  Stored in tests/golden/accuracy_<name>.npz:
  sha256_clean / sha256_noisy  SHA-256 of the two float32 sinograms' bytes
                               (a device run checks it built the same data)
  rmse_f32, norms_f32          per iteration: RMSE(x - phantom), ||A^T r||
  rmse_f64, norms_f64          the same for the FP64 oracle
  x_f32_sub, x_f64_sub         final iterates subsampled every `sub` voxels
  meta                         JSON: workload, seeds, iterations, sub
"""

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "ccpi-reconstruction_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle  # noqa: E402
from ccpi.reconstruction import workloads as W  # noqa: E402

WORKLOADS = {
    "cone512": dict(geom=W.CONE_512, ellipsoids=30, seed=2, i0=1e4, iters=10, sub=8),
    "cone256x100": dict(geom=W.CONE_768_AT_256, ellipsoids=50, seed=5, i0=5e3, iters=100,
                        sub=4),
}


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def data(name):
    """(oracle geometry, phantom, noise-free sinogram, noisy sinogram)."""
    w = WORKLOADS[name]
    g = W.cone_geometry(**w["geom"])
    og = oracle.Geometry(g.angles, g.h_pixels, g.v_pixels, g.origin, g.voxel_size, g.shape,
                         cone=True, source=g.source, det_x=g.detector_x)
    ph = W.ellipsoid_phantom_host(g, w["ellipsoids"], seed=w["seed"], dens=(0.0, 0.05))
    clean = og.fp(ph)
    noisy = W.poisson_log_host(clean, w["i0"], seed=w["seed"])
    return og, ph, clean, noisy


def main(name):
    w = WORKLOADS[name]
    t0 = time.time()
    og, ph, clean, noisy = data(name)
    print(f"{name}: data in {time.time() - t0:.1f} s", flush=True)
    out = {"sha256_clean": sha(clean), "sha256_noisy": sha(noisy)}
    del clean
    sub = w["sub"]
    for tag, dt in (("f32", np.float32), ("f64", np.float64)):
        t0 = time.time()
        x, _, norms, rmse, _ = oracle.cgls_track(og, noisy, w["iters"], ph, dtype=dt)
        print(f"{name}: {tag} CGLS {w['iters']} it in {time.time() - t0:.1f} s, "
              f"rmse {rmse[-1]:.6e}", flush=True)
        out[f"rmse_{tag}"] = rmse
        out[f"norms_{tag}"] = norms
        out[f"x_{tag}_sub"] = np.ascontiguousarray(x[::sub, ::sub, ::sub]).astype(np.float32)
    meta = dict(workload=name, geometry=w["geom"], ellipsoids=w["ellipsoids"], seed=w["seed"],
                i0=w["i0"], iterations=w["iters"], sub=sub,
                oracle_threads=os.environ.get("OMP_NUM_THREADS", str(os.cpu_count())))
    out["meta"] = np.array(json.dumps(meta))
    np.savez(os.path.join(HERE, f"accuracy_{name}.npz"), **out)
    print(f"{name}: written", flush=True)


if __name__ == "__main__":
    for n in sys.argv[1:] or list(WORKLOADS):
        main(n)
