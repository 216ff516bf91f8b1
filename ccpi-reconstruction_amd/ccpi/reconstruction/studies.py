"""Workload studies run on the device: precision of the CGLS reductions at
the 768^3 scale, and the differentiable TV-regularised reconstruction.

  python -m ccpi.reconstruction.studies precision [iterations]
  python -m ccpi.reconstruction.studies tv [steps]
  python -m ccpi.reconstruction.studies run <workload> [iterations] [--copy-calibration]

Each prints one JSON object per line (per-iteration records, then a
summary).
"""

from __future__ import annotations

import json
import math
import sys
import time

from . import workloads as W
from .autograd import data_fidelity, tv_isotropic
from .projectors import CglsSession, forward_project, projector_nnz


def _emit(obj):
    print(json.dumps(obj), flush=True)


def precision(iters: int = 30):
    """768^3 volume (1/3 mm voxels), 1200 projections over 360 deg, 1024^2
    detector (0.75 mm), SOD 500, SDD 1000; 50-ellipsoid phantom (seed 5),
    Poisson noise I0 = 5e3.  CGLS with FP64 reductions and with FP32
    reductions, RMSE vs the phantom and ||A^T r|| per iteration."""
    import torch
    g = W.cone_geometry(**W.CONE_768)
    ph = W.ellipsoid_phantom(g, 50, seed=5, dens=(0.0, 0.05))
    b = W.poisson_log(forward_project(g, ph), 5e3, seed=5)
    out = {}
    for name, fp32 in (("fp64-reductions", False), ("fp32-reductions", True)):
        x = torch.zeros(g.shape, dtype=torch.float32, device=b.device)
        s = CglsSession(g, b.clone(), x, iters, reduce_fp32=fp32)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rmse = []
        for it in range(iters):
            s.iterate(1)
            rmse.append(float(torch.sqrt(torch.mean((x - ph) ** 2))))
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        norms = s.norms()
        s.close()
        for it in range(iters):
            _emit({"variant": name, "iter": it + 1, "rmse": rmse[it], "norm_atr": norms[it]})
        out[name] = {"rmse_final": rmse[-1], "norm_final": float(norms[-1]),
                     "iter_s": iters / secs, "x": x.cpu()}
    xa, xb = out["fp64-reductions"].pop("x"), out["fp32-reductions"].pop("x")
    diff = float(torch.linalg.vector_norm(xa - xb) / torch.linalg.vector_norm(xa))
    _emit({"summary": "precision", "workload": "cone CGLS 768^3, 1200 proj, 1024^2",
           "iterations": iters, "variants": out, "rel_diff_fp32_vs_fp64_reductions": diff})


def tv(steps: int = 30, lam: float = 1e-3):
    """256^3 volume (1 mm voxels), 360 projections over 360 deg, 384^2
    detector (2 mm), SOD 500, SDD 1000; loss 0.5||Ax - b||^2 + lam TV(x)
    minimised by gradient descent through ForwardProjection (gradient of the
    data term = hand-written backprojection).  Step 1/L with L a power-
    iteration estimate of ||A^T A||."""
    import torch
    g = W.cone_geometry(**W.CONE_256)
    ph = W.ellipsoid_phantom(g, 30, seed=4, dens=(0.0, 0.05))
    b = W.poisson_log(forward_project(g, ph), 1e4, seed=4)
    from .projectors import backward_project
    v = torch.rand(g.shape, device=b.device, generator=torch.Generator(device=b.device).manual_seed(4))
    for _ in range(8):
        w = backward_project(g, forward_project(g, v))
        lip = float(torch.linalg.vector_norm(w) / torch.linalg.vector_norm(v))
        v = w / torch.linalg.vector_norm(w)
    step = 1.0 / (lip + 8.0 * lam)
    x = torch.zeros(g.shape, dtype=torch.float32, device=b.device, requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for it in range(steps):
        loss = data_fidelity(x, g, b) + lam * tv_isotropic(x)
        grad, = torch.autograd.grad(loss, x)
        with torch.no_grad():
            x -= step * grad.float()
            x.clamp_(min=0.0)
        rmse = float(torch.sqrt(torch.mean((x.detach() - ph) ** 2)))
        _emit({"step": it + 1, "loss": float(loss), "rmse": rmse})
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    _emit({"summary": "tv", "workload": "cone TV-GD 256^3, 360 proj, 384^2",
           "steps": steps, "lambda": lam, "lipschitz": lip, "steps_per_s": steps / secs,
           "nnz": projector_nnz(g)})


def run(workload: str, iters: int = 5, calibrate: bool = False):
    """A few CGLS iterations of one workload, for kernel traces and counter
    runs (rocprofv3) that should see one shape per kernel:
      cone512   the 512^3 cone scan (720 x 768^2);
      wide      the same scan through the XTek volume rule (768^3 of
                0.811 mm, the phased BP);
      cfg1      parallel 256 x 256 x 16, 180 angles, 367 x 16 detector;
      batch64   the 64 uneven parallel problems, plain (one at a time);
      batch64b  the same problems, batched (one persistent launch per step).
    ``calibrate`` first copies 1 GiB device to device (the known-size kernel
    that scales FETCH_SIZE / WRITE_SIZE counters)."""
    import numpy as np
    import torch

    from . import instruments as ins
    from .projectors import cgls, cgls_batched
    if calibrate:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        torch.empty_like(src).copy_(src)
        torch.cuda.synchronize()
        del src
    if workload in ("cone512", "wide", "cfg1"):
        if workload == "cone512":
            g = W.cone_geometry(**W.CONE_512)
        elif workload == "wide":
            ang = np.linspace(0, 360, 720, endpoint=False).astype(np.float32)
            g = ins.xtek_setup((720, 768, 768), ang, 1, -500.0, 500.0, 1.0, 1.0, 128.0).geometry
        else:
            g = W.parallel_geometry(256, 16, 180, 367)
        gen = torch.Generator(device="cuda").manual_seed(7)
        b = torch.rand(g.sino_shape, device="cuda", generator=gen)
        cgls(g, b, iters)
    elif workload in ("batch64", "batch64b"):
        geoms = W.batched_problems(64, seed=3)
        gen = torch.Generator(device="cuda").manual_seed(7)
        bs = [torch.rand(g.sino_shape, device="cuda", generator=gen) for g in geoms]
        if workload == "batch64b":
            cgls_batched(geoms, bs, iters)
        else:
            for g, b in zip(geoms, bs):
                cgls(g, b, iters)
    else:
        raise SystemExit(f"unknown workload {workload}")
    torch.cuda.synchronize()
    _emit({"summary": "run", "workload": workload, "iterations": iters})


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "precision"
    if which == "run":
        run(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else 5,
            calibrate="--copy-calibration" in sys.argv)
        sys.exit(0)
    n = int(sys.argv[2]) if len(sys.argv) > 2 else 30
    if which == "precision":
        precision(n)
    elif which == "tv":
        tv(n)
    else:
        raise SystemExit("usage: python -m ccpi.reconstruction.studies "
                         "precision|tv [n] | run <workload> [iterations]")
