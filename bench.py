#!/usr/bin/env python3
"""Headline benchmark: cone-beam CGLS on a 512^3 volume.

Workload: 512^3 float32 volume of 0.5 mm voxels (256 mm cube centred on the
rotation axis), 720 projections over 360 deg, flat 768 x 768 detector of
1.0 mm pixels, source-object 500 mm, source-detector 1000 mm; seeded
30-ellipsoid phantom (densities U[0, 0.05] /mm, seed 2) built on the host,
its projection on the device, Poisson noise at I0 = 1e4 drawn on the host
(numpy PCG64, seed 2) on the log-transformed sinogram.

A step is one CGLS iteration (one forward projection, one backprojection,
FP64 alpha/beta reductions and the fused vector updates) with every input
already resident in HBM.  `value` is CGLS iterations per second.

roofline: the dominant projector kernel's algorithmic bytes per launch,
4 * nnz(A) + 4 * (N_rays + N_vox) with nnz counted exactly on the device,
divided by its mean launch time from HIP events recorded on the launch
stream inside the timed iterations, against 8 TB/s.  `traffic` is the HBM-
side bytes per launch from the committed counter summary, reported only
while the summary's source stamp matches the kernel sources' code.

accuracy: 10 CGLS iterations on the same data as the committed FP64 / float
oracle runs (tests/golden/accuracy_cone512.npz, identical sinogram checked
by SHA-256): RMSE against the phantom of the device, of the FP64 oracle and
of the reference's float arithmetic.

cpu_baseline: the CPU oracle (a C restatement of the reference's f2D / b2D
and CGLS body, built with the reference's flags) runs one full CGLS
iteration of this workload on the host (forward and back projection of all
720 angles and the vector updates, timed inside the oracle).
"""

from __future__ import annotations

import argparse
import hashlib
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "ccpi-reconstruction_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
KERNEL_SOURCES = ("cone.hip", "parallel.hip", "par_body.hpp", "common.hpp")
# the sources each counter summary's kernels are built from
SUMMARY_SOURCES = {
    "cone512_projectors_counters.md": ("cone.hip", "common.hpp"),
    "parallel_256x256x16_cgls.md": ("parallel.hip", "par_body.hpp", "common.hpp"),
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def _code_text(text: str) -> str:
    """A C++ source without its comments, indentation and blank lines: the
    part of it that a rebuilt kernel depends on (no string literal in these
    sources holds a comment marker)."""
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    lines = (" ".join(line.split()) for line in text.splitlines())
    return "\n".join(line for line in lines if line)


def source_stamp(names=KERNEL_SOURCES) -> str:
    """SHA-256 (first 16 hex digits) of the code of projector kernel sources
    (by default all of them), comments and layout left out."""
    h = hashlib.sha256()
    for name in names:
        with open(os.path.join(ROOT, "ccpi-reconstruction_amd", "csrc", name)) as f:
            h.update(_code_text(f.read()).encode())
    return h.hexdigest()[:16]


def measured_traffic(kernel: str, summary: str = "cone512_projectors_counters.md"):
    """HBM-side bytes per launch of `kernel` from a committed counter summary
    under profiles/ (FETCH_SIZE and WRITE_SIZE runs scaled by a 1 GiB copy):
    one `TRAFFIC <kernel> <bytes>` line per kernel and a `SOURCES <stamp>`
    line.  Returns (bytes or None, "current" / "stale" / "missing")."""
    path = os.path.join(ROOT, "profiles", summary)
    val, stamp = None, None
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "TRAFFIC" and parts[1] == kernel:
                    val = int(parts[2])
                if len(parts) == 2 and parts[0] == "SOURCES":
                    stamp = parts[1]
    except OSError:
        return None, "missing"
    if val is None:
        return None, "missing"
    if stamp != source_stamp(SUMMARY_SOURCES.get(summary, KERNEL_SOURCES)):
        return None, "stale"
    return val, "current"


def sha256(a) -> str:
    import numpy as np
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def headline_data(geom, dev):
    """Phantom (host), its projection (device), Poisson noise (host): the
    data of the committed oracle runs.  Returns the phantom (host, device),
    the noisy sinogram (host, device) and the two sinograms' SHA-256."""
    import torch

    from ccpi.reconstruction import workloads as W
    from ccpi.reconstruction.projectors import forward_project
    ph = W.ellipsoid_phantom_host(geom, 30, seed=2, dens=(0.0, 0.05))
    ph_dev = torch.from_numpy(ph).to(dev)
    clean = forward_project(geom, ph_dev).cpu().numpy()   # also builds the plan
    sha_clean = sha256(clean)
    noisy = W.poisson_log_host(clean, 1e4, seed=2)
    del clean
    return ph, ph_dev, noisy, torch.from_numpy(noisy).to(dev), sha_clean, sha256(noisy)


def accuracy_512(geom, b, phantom, sha_clean, sha_noisy, iters=10):
    """CGLS from x = 0 on the headline data against the committed oracle
    runs of the same data."""
    import numpy as np
    import torch

    from ccpi.reconstruction.projectors import CglsSession
    path = os.path.join(ROOT, "tests", "golden", "accuracy_cone512.npz")
    out = {"iterations": iters}
    x = torch.zeros(geom.shape, dtype=torch.float32, device=b.device)
    s = CglsSession(geom, b.clone(), x, iters)
    rmse = []
    for _ in range(iters):
        s.iterate(1)
        rmse.append(float(torch.sqrt(torch.mean((x - phantom).double() ** 2))))
    norms = s.norms()
    s.close()
    out["rmse_gpu"] = rmse[-1]
    out["norm_atr_gpu"] = float(norms[-1])
    if not os.path.exists(path):
        out["oracle"] = "fixture missing"
        return out
    f = np.load(path)
    n = min(iters, len(f["rmse_f64"]))
    same = sha_clean == str(f["sha256_clean"]) and sha_noisy == str(f["sha256_noisy"])
    out.update({
        "same_data_as_oracle_runs": same,
        "rmse_fp64_oracle": float(f["rmse_f64"][n - 1]),
        "rmse_ref_float": float(f["rmse_f32"][n - 1]),
        "norm_atr_fp64_oracle": float(f["norms_f64"][n - 1]),
        "norm_atr_ref_float": float(f["norms_f32"][n - 1]),
    })
    out["rmse_rel_err_vs_fp64"] = abs(rmse[n - 1] / out["rmse_fp64_oracle"] - 1)
    out["ref_float_rmse_rel_err_vs_fp64"] = abs(out["rmse_ref_float"] /
                                                out["rmse_fp64_oracle"] - 1)
    sub = int(json.loads(str(f["meta"]))["sub"])
    xs = x[::sub, ::sub, ::sub].cpu().numpy().astype(np.float64)
    x64 = f["x_f64_sub"].astype(np.float64)
    nrm = float(np.linalg.norm(x64))
    out["x_rel_dist_vs_fp64_subsampled"] = float(np.linalg.norm(xs - x64)) / nrm
    out["ref_float_x_rel_dist_vs_fp64_subsampled"] = float(
        np.linalg.norm(f["x_f32_sub"].astype(np.float64) - x64)) / nrm
    return out


def cpu_baseline(geom, noisy_np):
    """One full CGLS iteration of the workload by the oracle on the host
    (all angles; forward projection, backprojection and vector updates
    timed inside the oracle; the initial A^T b is not counted)."""
    import oracle
    og = oracle.Geometry(geom.angles, geom.h_pixels, geom.v_pixels, geom.origin,
                         geom.voxel_size, geom.shape, cone=True, source=geom.source,
                         det_x=geom.detector_x)
    t0 = time.perf_counter()
    _, _, _, _, secs = oracle.cgls_track(og, noisy_np, 1)
    wall = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": 1.0 / float(secs[0]),
        "unit": "iter/s",
        "cores": cores,
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": (f"one full CGLS iteration of the 512^3 workload (FP + BP of all 720 "
                   f"angles, float, OpenMP, and the vector updates): {secs[0]:.1f} s; "
                   f"{wall:.1f} s with the initial A^T b"),
    }


def batched_vs_plain(iters: int, small: bool = False):
    """Batched parallel-beam CGLS: 64 uneven problems, batched (one
    persistent launch per step, graph replay) vs the same problems one after
    another through the one-problem path.  Per-iteration time is
    (T(iters) - T(0)) / iters for both, so set-up is excluded alike.
    ``small``: the same comparison on 256 problems too small to fill the
    GPU one at a time (N in {64, 128}, S <= 4)."""
    import numpy as np
    import torch

    from ccpi.reconstruction import workloads as W
    from ccpi.reconstruction.projectors import (CglsSession, cgls_batched, forward_project,
                                                projector_nnz)

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
        sess.append((CglsSession(g, b.clone(), x, iters + 1, timing=True), x))
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
    # the plain path's projector launches (HIP events of every session) and
    # the mix's algorithmic bytes: its projector throughput
    fp_ms = bp_ms = 0.0
    n_fp = n_bp = 0
    bytes_alg = 0
    for (ss, _), g in zip(sess, geoms):
        f, nf, bk, nb = ss.timings()
        fp_ms += f
        bp_ms += bk
        n_fp += nf
        n_bp += nb
        nnz = projector_nnz(g)
        bytes_alg += 4 * nnz + 4 * (int(np.prod(g.shape)) + int(np.prod(g.sino_shape)))
    launches = len(geoms)
    fp_it = fp_ms / max(n_fp // launches, 1)
    bp_it = bp_ms / max(n_bp // launches, 1)
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
        "plain_projectors": {
            "fp_ms_per_iter": fp_it, "bp_ms_per_iter": bp_it, "bytes_alg": bytes_alg,
            "fp_frac": bytes_alg / (fp_it * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "bp_frac": bytes_alg / (bp_it * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
    }


def parallel_cfg1(cpu: bool):
    """Plain parallel-beam CGLS, 256 x 256 x 16 slices (1 mm), 180 angles
    over [0, 180 deg), 367 x 16 detector, 10-ellipsoid phantom (seed 1,
    densities U[0.1, 1]), noise-free sinogram, 20 iterations from x0 = 0;
    the CPU oracle runs the same 20 iterations.  The projectors' launch
    times come from HIP events of a timed CGLS session (roofline)."""
    import numpy as np
    import torch

    from ccpi.reconstruction import workloads as W
    from ccpi.reconstruction.projectors import (CglsSession, cgls, forward_project,
                                                projector_nnz)

    g = W.parallel_geometry(256, 16, 180, 367)
    ph = torch.from_numpy(W.ellipsoid_phantom_host(g, 10, seed=1, dens=(0.1, 1.0))).cuda()
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
    nnz = projector_nnz(g)
    bytes_alg = 4 * nnz + 4 * (int(np.prod(g.shape)) + int(np.prod(g.sino_shape)))
    xs = torch.zeros(g.shape, dtype=torch.float32, device="cuda")
    s = CglsSession(g, b.clone(), xs, 20, timing=True)
    s.iterate(20)
    fp_ms, nf, bp_ms, nb = s.timings()
    s.close()
    fp_ms, bp_ms = fp_ms / max(nf, 1), bp_ms / max(nb, 1)
    dom, dom_ms = ("par_bp_lds_kernel", bp_ms) if bp_ms >= fp_ms else ("par_fp_kernel", fp_ms)
    ach = bytes_alg / (dom_ms * 1e-3) / 1e9
    traffic, traffic_state = measured_traffic(dom, "parallel_256x256x16_cgls.md")
    out = {"workload": "parallel-beam CGLS 256x256x16, 180 angles, 367x16 detector, "
                       "20 iterations, noise-free",
           "iter_s": 20.0 / secs, "ms_per_iter": secs / 20 * 1e3,
           "rmse_vs_phantom": rmse, "nnz": nnz, "fp_ms": fp_ms, "bp_ms": bp_ms,
           "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                        "traffic_source": traffic_state, "bytes_alg": bytes_alg,
                        "fp_frac": bytes_alg / (fp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "bp_frac": bytes_alg / (bp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}}
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
    """Cone-beam SIRT and MLEM at the 512^3 geometry, `iters` iterations
    each on the headline sinogram; per-iteration time is
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


def errors_vs_fp64(iters: int = 10):
    """Error against the FP64 oracle of every iteration body at the 512^3
    scan's geometry scaled to 128^3 (1 mm voxels, 180 projections, 192^2
    detector of 2 mm, SOD 500, SDD 1000), 30-ellipsoid phantom (seed 2),
    Poisson noise I0 = 1e4: relative l2 distance of the device result and
    of the oracle's float result (the reference's arithmetic) to the FP64
    oracle's, after `iters` iterations."""
    import numpy as np
    import torch

    import oracle
    from ccpi.reconstruction import workloads as W
    from ccpi.reconstruction.projectors import (TIKHONOV, TV, cgls, cgls_reg, forward_project,
                                                mlem, sirt)
    g = W.cone_geometry(128, 1.0, 180, 192, 2.0, 500.0, 1000.0)
    og = oracle.Geometry(g.angles, g.h_pixels, g.v_pixels, g.origin, g.voxel_size, g.shape,
                         cone=True, source=g.source, det_x=g.detector_x)
    ph = W.ellipsoid_phantom_host(g, 30, seed=2, dens=(0.0, 0.05))
    clean = forward_project(g, torch.from_numpy(ph).cuda()).cpu().numpy()
    bn = W.poisson_log_host(clean, 1e4, seed=2)
    b = torch.from_numpy(bn).cuda()
    runs = {
        "cgls": (lambda: cgls(g, b, iters)[0],
                 lambda dt: oracle.cgls(og, bn.astype(dt), iters, dtype=dt)[0]),
        "sirt": (lambda: sirt(g, b, iters),
                 lambda dt: oracle.sirt(og, bn.astype(dt), iters, dtype=dt)),
        "mlem": (lambda: mlem(g, b, iters),
                 lambda dt: oracle.mlem(og, bn.astype(dt), iters, dtype=dt)),
        "mlem_accum_fp64": (lambda: mlem(g, b, iters, accum_fp64=True),
                            lambda dt: oracle.mlem(og, bn.astype(dt), iters, dtype=dt)),
        "cgls_tikhonov": (lambda: cgls_reg(g, b, iters, TIKHONOV, 0.1)[0],
                          lambda dt: oracle.cgls_reg(og, bn.astype(dt), iters, oracle.TIKHONOV,
                                                     0.1, dtype=dt)[0]),
        "cgls_tv": (lambda: cgls_reg(g, b, iters, TV, 1e-4)[0],
                    lambda dt: oracle.cgls_reg(og, bn.astype(dt), iters, oracle.TV, 1e-4,
                                               dtype=dt)[0]),
    }
    out = {"workload": "cone 128^3 (1 mm), 180 proj / 360 deg, 192^2 detector (2 mm), "
                       f"SOD 500, SDD 1000, Poisson I0 1e4, {iters} iterations"}
    for name, (dev_fn, or_fn) in runs.items():
        xd = dev_fn().cpu().numpy().astype(np.float64)
        x64 = or_fn(np.float64)
        x32 = or_fn(np.float32).astype(np.float64)
        nrm = float(np.linalg.norm(x64))
        out[name] = {"err_vs_fp64": float(np.linalg.norm(xd - x64)) / nrm,
                     "ref_float_err_vs_fp64": float(np.linalg.norm(x32 - x64)) / nrm}
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
    ap.add_argument("--cpu", type=int, default=1,
                    help="run the CPU oracle legs (baseline iteration, checkers)")
    ap.add_argument("--batched-iters", type=int, default=20,
                    help="CGLS iterations of the batched-vs-plain comparison (0 skips it)")
    ap.add_argument("--extras", type=int, default=1,
                    help="also measure the parallel CGLS, SIRT/MLEM and regularised bodies")
    ap.add_argument("--wide", type=int, default=1,
                    help="with --extras: also time the wide-cone BP (0 disables it)")
    ap.add_argument("--accuracy", type=int, default=1,
                    help="run the 10-iteration accuracy comparison with the oracle runs")
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
    from ccpi.reconstruction.projectors import CglsSession, projector_nnz

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    cfg = W.CONE_512
    geom = W.cone_geometry(**cfg)
    log("bench: phantom (host) + sinogram (device) + Poisson noise (host)")
    ph_np, phantom, noisy_np, b, sha_clean, sha_noisy = headline_data(geom, dev)
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
    del x

    accuracy = None
    if args.accuracy:
        log("bench: accuracy against the oracle runs (10 iterations)")
        accuracy = accuracy_512(geom, b, phantom, sha_clean, sha_noisy)

    batched = None
    if args.batched_iters > 0:
        log("bench: batched vs plain (64 parallel-beam problems)")
        batched = batched_vs_plain(args.batched_iters)

    extras = None
    if args.extras:
        log("bench: parallel CGLS, cone SIRT / MLEM, regularised CGLS")
        extras = {"cfg1": parallel_cfg1(cpu=args.cpu > 0),
                  "cfg2": cone_sirt_mlem(geom, b, 4),
                  "cgls_reg": cone_cgls_reg(geom, b, 3)}
        if args.cpu > 0:
            log("bench: errors against the FP64 oracle (128^3)")
            extras["err_vs_fp64"] = errors_vs_fp64()
        if args.batched_iters > 0:
            log("bench: batched vs plain on small problems")
            extras["batched_small"] = batched_vs_plain(args.batched_iters, small=True)
        if args.wide:
            log("bench: wide-cone BP (XTek volume rule)")
            extras["cone_wide_bp"] = cone_wide_bp()

    base = None
    if args.cpu > 0:
        log("bench: CPU oracle, one full CGLS iteration")
        del b, phantom
        torch.cuda.empty_cache()
        base = cpu_baseline(geom, noisy_np)

    traffic, traffic_state = measured_traffic(dom_name)
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
            "traffic": traffic,
            "traffic_source": traffic_state,
            "bytes_alg": bytes_alg,
            "nnz": nnz,
        },
        "fp_ms": fp_ms,
        "bp_ms": bp_ms,
        "fp_gbs": bytes_alg / (fp_ms * 1e-3) / 1e9,
        "bp_gbs": bytes_alg / (bp_ms * 1e-3) / 1e9,
        "rmse_vs_phantom": rmse,
        "norm_atr_last": float(norms[-1]) if len(norms) else None,
        "accuracy": accuracy,
        "batched": batched,
        "extras": extras,
        "cpu_baseline": base,
        "kernel_sources": source_stamp(),
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
