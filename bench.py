#!/usr/bin/env python3
"""Headline benchmark: cone-beam CGLS on a 512^3 volume.

Workload: 512^3 float32 volume of 0.5 mm voxels (256 mm cube centred on the
rotation axis), 720 projections over 360 deg, flat 768 x 768 detector of
1.0 mm pixels, source-object 500 mm, source-detector 1000 mm; seeded
30-ellipsoid phantom (densities U[0, 0.05] /mm, seed 2), Poisson noise at
I0 = 1e4 on the log-transformed sinogram.

A step is one CGLS iteration (one forward projection, one backprojection,
FP64 alpha/beta reductions and the fused vector updates) with every input
already resident in HBM.  `value` is CGLS iterations per second.

roofline: the dominant projector kernel's algorithmic bytes per launch,
4 * nnz(A) + 4 * (N_rays + N_vox) with nnz counted exactly on the device,
divided by its mean launch time from HIP events recorded on the launch
stream inside the timed iterations, against 8 TB/s.

cpu_baseline: the CPU oracle (a C restatement of the reference's f2D / b2D,
built with the reference's flags) timed on this host on a sample of the
projection angles, scaled to one CGLS iteration.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "ccpi-reconstruction_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(kernel: str):
    """HBM-side bytes per launch of `kernel` from the committed counter
    summary (FETCH_SIZE and WRITE_SIZE runs scaled by a 1 GiB copy); the
    summary holds one `TRAFFIC <kernel> <bytes>` line per kernel."""
    path = os.path.join(ROOT, "profiles", "cone512_projectors_counters.md")
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "TRAFFIC" and parts[1] == kernel:
                    return int(parts[2])
    except OSError:
        pass
    return None


def cpu_baseline(geom, phantom_np, n_sample: int):
    """Time the oracle's cone f2D + b2D on `n_sample` evenly spaced angles
    of the workload, scaled to all angles = one CGLS iteration."""
    import numpy as np

    import oracle
    sel = np.linspace(0, geom.angles.size, n_sample, endpoint=False).astype(int)
    og = oracle.Geometry(geom.angles[sel], geom.h_pixels, geom.v_pixels, geom.origin,
                         geom.voxel_size, geom.shape, cone=True, source=geom.source,
                         det_x=geom.detector_x)
    t0 = time.perf_counter()
    y = og.fp(phantom_np)
    t1 = time.perf_counter()
    og.bp(y)
    t2 = time.perf_counter()
    scale = geom.angles.size / n_sample
    iter_s = (t1 - t0 + t2 - t1) * scale
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": 1.0 / iter_s,
        "unit": "iter/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"oracle cone f2D+b2D (float, OpenMP) on {n_sample} of "
                   f"{geom.angles.size} angles of the 512^3 workload, "
                   f"FP {t1 - t0:.2f} s + BP {t2 - t1:.2f} s, scaled x{scale:.0f}"),
    }


def batched_vs_plain(iters: int, small: bool = False):
    """Config 3: 64 uneven parallel-beam CGLS problems, batched (one
    persistent launch per step, graph replay) vs the same problems one after
    another through the one-problem path.  Per-iteration time is
    (T(iters) - T(0)) / iters for both, so set-up is excluded alike.
    ``small``: the same comparison on 256 problems too small to fill the
    GPU one at a time (N in {64, 128}, S <= 4)."""
    import numpy as np
    import torch

    from ccpi.reconstruction import workloads as W
    from ccpi.reconstruction.projectors import CglsSession, cgls_batched, forward_project

    geoms = (W.batched_problems(256, seed=5, sizes=(64, 128), slices=(1, 4), angles=(90, 180))
             if small else W.batched_problems(64, seed=3))
    phs, bs = [], []
    for i, g in enumerate(geoms):
        ph = W.ellipsoid_phantom(g, 6, seed=300 + i, dens=(0.1, 1.0))
        b = W.gaussian_noise(forward_project(g, ph), 0.01, seed=300 + i)
        phs.append(ph)
        bs.append(b)

    def run_batched(n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xs, _, _ = cgls_batched(geoms, bs, n)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, xs

    run_batched(1)                       # autotune + plans
    t0b, _ = run_batched(0)
    tb, xb = run_batched(iters)
    per_b = (tb - t0b) / iters

    sess = []
    for g, b in zip(geoms, bs):
        x = torch.zeros(g.shape, dtype=torch.float32, device=b.device)
        sess.append((CglsSession(g, b.clone(), x, iters + 1), x))
    torch.cuda.synchronize()
    sess[0][0].iterate(1)                # warm the one-problem kernels
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for ss, _ in sess[1:]:
        ss.iterate(iters)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sess[0][0].iterate(iters)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    per_p = (t1 - t0 + t2 - t1) / iters
    rm_b, rm_p = [], []
    for ph, x1, (ss, x2) in zip(phs, xb, sess):
        rm_b.append(float(torch.sqrt(torch.mean((x1 - ph) ** 2))))
    # the first session ran one extra iteration; compare the rest
    for ph, (ss, x2) in zip(phs[1:], sess[1:]):
        rm_p.append(float(torch.sqrt(torch.mean((x2 - ph) ** 2))))
    rel = max(abs(a - b) / max(b, 1e-30) for a, b in zip(rm_b[1:], rm_p))
    for ss, _ in sess:
        ss.close()
    return {
        "workload": ("256 small parallel-beam CGLS problems, N in {64, 128}, S in [1, 4], "
                     "{90,180} angles, width ceil(sqrt2 N), 1% Gaussian noise, seed 5"
                     if small else
                     "64 parallel-beam CGLS problems, N in {128..512}, S in [1, 32], "
                     "{90,180,360} angles, width ceil(sqrt2 N), 1% Gaussian noise, seed 3"),
        "iterations": iters,
        "batched_iter_s": 1.0 / per_b,
        "plain_iter_s": 1.0 / per_p,
        "speedup": per_p / per_b,
        "rmse_mean": float(np.mean(rm_b)),
        "rmse_max_rel_diff_batched_vs_plain": rel,
    }


def parallel_cfg1(cpu: bool):
    """Config 1: plain parallel-beam CGLS, 256 x 256 x 16 slices (1 mm),
    180 angles over [0, 180 deg), 367 x 16 detector, 10-ellipsoid phantom
    (seed 1, densities U[0.1, 1]), noise-free sinogram, 20 iterations from
    x0 = 0; the CPU oracle runs the same 20 iterations."""
    import numpy as np
    import torch

    from ccpi.reconstruction import workloads as W
    from ccpi.reconstruction.projectors import cgls, forward_project, projector_nnz

    g = W.parallel_geometry(256, 16, 180, 367)
    ph = W.ellipsoid_phantom(g, 10, seed=1, dens=(0.1, 1.0))
    b = forward_project(g, ph)
    cgls(g, b, 1)                                # plan + warm kernels
    torch.cuda.synchronize()
    reps = []
    for _ in range(3):
        t0 = time.perf_counter()
        x, _, _ = cgls(g, b, 20)
        torch.cuda.synchronize()
        reps.append(time.perf_counter() - t0)
    secs = sorted(reps)[1]
    rmse = float(torch.sqrt(torch.mean((x - ph) ** 2)))
    out = {"workload": "parallel-beam CGLS 256x256x16, 180 angles, 367x16 detector, "
                       "20 iterations, noise-free",
           "iter_s": 20.0 / secs, "ms_per_iter": secs / 20 * 1e3,
           "rmse_vs_phantom": rmse, "nnz": projector_nnz(g)}
    if cpu:
        import oracle
        og = oracle.Geometry(g.angles, g.h_pixels, g.v_pixels, g.origin, g.voxel_size,
                             g.shape)
        bn = b.cpu().numpy()
        t0 = time.perf_counter()
        xo, _, _ = oracle.cgls(og, bn, 20)
        tc = time.perf_counter() - t0
        out["cpu_oracle_iter_s"] = 20.0 / tc
        out["cpu_cores"] = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        xg = x.cpu().numpy()
        out["max_abs_diff_vs_cpu_oracle"] = float(np.abs(xg - xo).max())
        # error against the FP64 oracle of the same model (relative l2)
        x64, _, _ = oracle.cgls(og, bn.astype(np.float64), 20, dtype=np.float64)
        nrm = float(np.linalg.norm(x64))
        out["rel_err_vs_fp64_oracle"] = float(np.linalg.norm(xg - x64)) / nrm
        out["cpu_float_oracle_rel_err_vs_fp64"] = float(np.linalg.norm(xo - x64)) / nrm
    return out


def cone_cgls_reg(geom, b, iters: int):
    """Tikhonov- and TV-regularised CGLS (cgls.cpp:343-435) at the 512^3
    geometry on the headline sinogram; per-iteration time as in
    cone_sirt_mlem."""
    import torch

    from ccpi.reconstruction.projectors import TIKHONOV, TV, cgls_reg
    out = {}
    for name, kind, lam in (("tikhonov", TIKHONOV, 0.1), ("tv", TV, 1e-4)):
        ts = []
        for n in (1, iters):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            cgls_reg(geom, b, n, kind, lam)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        out[f"cgls_{name}_iter_s"] = (iters - 1) / max(ts[1] - ts[0], 1e-9)
    out["workload"] = f"cone-beam regularised CGLS 512^3, {iters} iterations each"
    return out


def cone_sirt_mlem(geom, b, iters: int):
    """Config 2: cone-beam SIRT and MLEM at the 512^3 geometry, `iters`
    iterations each on the headline sinogram; per-iteration time is
    (T(iters) - T(1)) / (iters - 1), so the set-up projections are excluded."""
    import torch

    from ccpi.reconstruction.projectors import mlem, sirt
    out = {}
    for name, fn in (("sirt", sirt), ("mlem", mlem)):
        ts = []
        for n in (1, iters):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn(geom, b, n)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        out[f"{name}_iter_s"] = (iters - 1) / max(ts[1] - ts[0], 1e-9)
    out["workload"] = f"cone-beam SIRT and MLEM 512^3, {iters} iterations each"
    return out


def cone_wide_bp(reps: int = 3):
    """The headline scan read through the XTek route (xtek_setup: 720 x 768^2
    detector of 1 mm pixels, SOD 500 mm, SDD 1000 mm), whose volume rule
    (xtek.cpp:760-798) gives 768^3 voxels of 0.811 mm with corners 60 mm
    from the source: up to 13.5 detector rows per slice, so the BP takes its
    phased form.  Best of `reps` BPs / FPs of a seeded random sinogram /
    volume (HIP events on the launch stream), and the BP's algorithmic
    bytes as for the roofline."""
    import numpy as np
    import torch

    from ccpi.reconstruction import instruments as ins
    from ccpi.reconstruction import projectors as P
    na, nh, nv = 720, 768, 768
    ang = np.linspace(0, 360, na, endpoint=False).astype(np.float32)
    g = ins.xtek_setup((na, nh, nv), ang, 1, -500.0, 500.0, 1.0, 1.0, 128.0).geometry
    gen = torch.Generator(device="cuda").manual_seed(0)
    s = torch.rand(g.sino_shape, device="cuda", generator=gen)
    x = torch.rand(g.shape, device="cuda", generator=gen)
    nnz = P.projector_nnz(g)

    def timed(f):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        f()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    vol = torch.zeros(g.shape, device="cuda")
    sino = torch.zeros(g.sino_shape, device="cuda")
    bps, fps = [], []
    for _ in range(reps):
        vol.zero_()
        bps.append(timed(lambda: P.backward_project(g, s, vol)))
        sino.zero_()
        fps.append(timed(lambda: P.forward_project(g, x, sino)))
    bp, fp = min(bps), min(fps)
    nvox = int(np.prod(g.shape))
    nray = int(np.prod(g.sino_shape))
    bytes_alg = 4 * nnz + 4 * (nray + nvox)
    gbs = bytes_alg / (bp * 1e-3) / 1e9
    return {"workload": "cone-beam BP, XTek volume rule: 768^3 of 0.811 mm, 720 x 768^2 "
                        "detector (1 mm), SOD 500 mm, SDD 1000 mm",
            "bp_ms": bp, "fp_ms": fp, "nnz": int(nnz), "bp_gbs": gbs,
            "bp_roofline_frac": gbs / HBM_PEAK_GBS}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=12,
                    help="angles in the CPU-oracle sample (0 disables it)")
    ap.add_argument("--batched-iters", type=int, default=20,
                    help="CGLS iterations of the batched-vs-plain comparison (0 skips it)")
    ap.add_argument("--extras", type=int, default=1,
                    help="also measure config 1 (parallel CGLS) and config 2 SIRT/MLEM")
    ap.add_argument("--wide", type=int, default=1,
                    help="with --extras: also time the wide-cone BP (0 disables it)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device to device (the known-size kernel that "
                         "scales FETCH_SIZE / WRITE_SIZE counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path, "
              "run it with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from ccpi.reconstruction import workloads as W
    from ccpi.reconstruction.projectors import (CglsSession, forward_project,
                                                projector_nnz)

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    cfg = W.CONE_512
    geom = W.cone_geometry(**cfg)
    log("bench: phantom + sinogram")
    phantom = W.ellipsoid_phantom(geom, 30, seed=2, dens=(0.0, 0.05), device=dev)
    clean = forward_project(geom, phantom)            # also builds the plan
    b = W.poisson_log(clean, 1e4, seed=2)
    del clean
    x = torch.zeros(geom.shape, dtype=torch.float32, device=dev)
    nnz = projector_nnz(geom)
    n_vox = int(np.prod(geom.shape))
    n_rays = int(np.prod(geom.sino_shape))
    bytes_alg = 4 * nnz + 4 * (n_rays + n_vox)

    # the session keeps its residual in the sinogram buffer it is given
    sess = CglsSession(geom, b.clone(), x, args.warmup + args.steps, timing=True)
    log(f"bench: {args.warmup} warmup iterations")
    sess.iterate(args.warmup)
    torch.cuda.synchronize()
    fp0, nf0, bp0, nb0 = sess.timings()
    log(f"bench: {args.steps} timed iterations")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sess.iterate(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    fp1, nf1, bp1, nb1 = sess.timings()
    norms = sess.norms()
    sess.close()
    secs = t1 - t0
    fp_ms = (fp1 - fp0) / max(nf1 - nf0, 1)
    bp_ms = (bp1 - bp0) / max(nb1 - nb0, 1)
    dom_name, dom_ms = ("cone_bp_kernel", bp_ms) if bp_ms >= fp_ms else ("cone_fp_kernel", fp_ms)
    achieved = bytes_alg / (dom_ms * 1e-3) / 1e9
    rmse = float(torch.sqrt(torch.mean((x - phantom) ** 2)))
    log(f"bench: {secs / args.steps * 1e3:.1f} ms/iter, FP {fp_ms:.1f} ms, BP {bp_ms:.1f} ms")

    batched = None
    if args.batched_iters > 0:
        log("bench: batched vs plain (64 parallel-beam problems)")
        batched = batched_vs_plain(args.batched_iters)

    extras = None
    if args.extras:
        log("bench: config 1 (parallel CGLS) and config 2 SIRT / MLEM")
        extras = {"cfg1": parallel_cfg1(cpu=args.cpu_sample > 0),
                  "cfg2": cone_sirt_mlem(geom, b, 4),
                  "cgls_reg": cone_cgls_reg(geom, b, 3)}
        if args.batched_iters > 0:
            log("bench: batched vs plain on small problems")
            extras["batched_small"] = batched_vs_plain(args.batched_iters, small=True)
        if args.wide:
            log("bench: wide-cone BP (XTek volume rule)")
            extras["cone_wide_bp"] = cone_wide_bp()

    base = None
    if args.cpu_sample > 0:
        log("bench: CPU oracle sample")
        base = cpu_baseline(geom, phantom.cpu().numpy(), args.cpu_sample)

    line = {
        "metric": "CGLS iter/s + projector HBM GB/s, cone-beam 512^3",
        "value": args.steps / secs,
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded 30-ellipsoid phantom, Poisson noise I0=1e4",
        "config": {
            "workload": "cone-beam CGLS 512^3 (0.5 mm voxels), 720 proj / 360 deg, "
                        "768x768 detector (1 mm), SOD 500 mm, SDD 1000 mm",
            "reductions": "fp64",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom_name),
            "bytes_alg": bytes_alg,
            "nnz": nnz,
        },
        "fp_ms": fp_ms,
        "bp_ms": bp_ms,
        "fp_gbs": bytes_alg / (fp_ms * 1e-3) / 1e9,
        "bp_gbs": bytes_alg / (bp_ms * 1e-3) / 1e9,
        "rmse_vs_phantom": rmse,
        "norm_atr_last": float(norms[-1]) if len(norms) else None,
        "batched": batched,
        "extras": extras,
        "cpu_baseline": base,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
