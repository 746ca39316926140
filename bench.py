"""Benchmark: geodesics/s for Kerr a=0.9 at 1024x1024 on one MI355X.

A step is one render of the whole 1024x1024 jittered Kerr a=0.9 thin-disk
scene (camera r=100M, inclination 60 deg, FOV 25 deg, rtol 1e-8, atol 1e-10)
by the mixed FP32/FP64 kernel, with the per-ray outputs written to device
memory.  The plain FP64 kernel is timed the same way beside it, and the
mixed result is compared with the FP64 one (class agreement, final-state
error).  The roofline object reports algorithmic FLOPs of the trace kernel
over its hipEvent time against the vector (VALU) peak; the CPU baseline is
the FP64 oracle on every ray of the same scene, whose outputs are also
compared with both GPU kernels.  The same line carries one timed frame of
the near-extremal 4096x4096 workload and of the 64-scene batch (FP64 and
mixed, roofline fractions, oracle on a stride sample), and the adjoint's
gradients/s on the 512x512 differentiable render with its error against
central differences of the oracle's loss.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--headline-only]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic FLOPs (add, sub, mul, div, sqrt, sin, cos, pow = 1 each; the
# FP64 kernels contract nothing) counted from the kernel source after common
# subexpressions: Kerr RHS 141 ops for the Christoffel symbols + 96 for the
# 32 non-zero contraction terms; Schwarzschild RHS 44; stage combinations,
# error norm and controller of one attempt that reaches the error estimate
# 318 (rtol/atol mode).
F_RHS = {"kerr": 237, "schwarzschild": 44}
F_COMB = 318
PEAK_FP64 = 78.6    # TFLOP/s, MI355X vector FP64 (published)
PEAK_FP32 = 157.3   # TFLOP/s, MI355X vector FP32 (published; needs packed math)


def profiled_traffic(profile: str, kernel: str):
    """HBM bytes per launch of `kernel` from a committed counter summary (FETCH_SIZE and WRITE_SIZE passes of
    rocprofv3, scaled by a 512 MiB copy), or None when the summary or the kernel is absent."""
    path = os.path.join(ROOT, "profiles", profile)
    try:
        col = None
        with open(path) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if cells and cells[0].startswith("derived"):
                    col = next((i for i, c in enumerate(cells) if c == f"`{kernel}`"), None)
                elif col is not None and cells[0] == "HBM traffic per launch (calibrated, MB)":
                    return float(cells[col]) * 1e6
    except (OSError, ValueError, IndexError):
        return None
    return None

METRIC = "geodesics/s Kerr a=0.9 1024²: mixed FP32 vs FP64 at ≤1e-6 err; adjoint grads/s"


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def _dev_outputs(torch, n):
    d = "cuda"
    return {
        "cls": torch.empty(n, dtype=torch.uint8, device=d), "r": torch.empty(n, dtype=torch.float64, device=d),
        "theta": torch.empty(n, dtype=torch.float64, device=d), "phi": torch.empty(n, dtype=torch.float64, device=d),
        "intensity": torch.empty(n, dtype=torch.float64, device=d),
        "attempts": torch.empty(n, dtype=torch.int32, device=d), "accepted": torch.empty(n, dtype=torch.int32, device=d),
    }


def _ptrs(o):
    return {k: v.data_ptr() for k, v in o.items()}


def time_precision(torch, rend, scene, precision, steps, warmup, outs):
    ptrs = _ptrs(outs)
    for _ in range(warmup):
        rend.render(scene, precision, outputs=dict(ptrs))
    torch.cuda.synchronize()
    kms, stats = [], None
    t0 = time.perf_counter()
    for _ in range(steps):
        o = rend.render(scene, precision, outputs=dict(ptrs))
        kms.append(o["stats"]["kernel_ms"])
        stats = o["stats"]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, kms, stats


def flops(stats, metric_kind):
    f_rhs = F_RHS[metric_kind]
    f32 = stats["rhs_evals_fp32"] * f_rhs
    f64 = (stats["rhs_evals"] - stats["rhs_evals_fp32"]) * f_rhs + stats["attempts_full"] * F_COMB
    return f32, f64


def cpu_baseline(scene, stride=1):
    """FP64 oracle on every `stride`-th row and column of the same rays (stride 1: all of them)."""
    from oracle import Oracle
    w = scene.width
    idx = np.array([py * w + px for py in range(0, scene.height, stride) for px in range(0, w, stride)],
                   dtype=np.uint64)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    ora = Oracle(threads=cores)
    t0 = time.perf_counter()
    ref = ora.render_rays(scene, idx, state=False)
    dt = time.perf_counter() - t0
    one = Oracle(threads=1)
    sub = idx[:: max(1, len(idx) // 2048)]
    t1 = time.perf_counter()
    one.render_rays(scene, sub, state=False)
    dt1 = time.perf_counter() - t1
    return {"value": len(idx) / dt, "unit": "geodesics/s", "cores": cores, "kind": "port",
            "sample": f"every {stride}. row and column of the {scene.width}x{scene.height} jittered Kerr a=0.9 scene "
                      f"({len(idx)} rays, {dt:.1f} s on {cores} threads); 1 thread: {len(sub) / dt1:.1f} geodesics/s "
                      f"on {len(sub)} rays",
            "value_1core": len(sub) / dt1}, ref


def com_drift(rend, scene, precision):
    """Drift of E, L_z, Carter Q from the camera state to the final state (untimed render with states)."""
    s0 = rend.camera_rays(scene)
    o = rend.render(scene, precision, state=True)
    s1 = np.ascontiguousarray(o["state"].T)
    c0 = rend.constants_of_motion(scene.metric, s0)
    c1 = rend.constants_of_motion(scene.metric, s1)
    d = np.abs(c1 - c0)
    rel = {"E": d[:, 0] / np.abs(c0[:, 0]), "L_z": d[:, 1] / (np.abs(c0[:, 1]) + 1e-3 * np.abs(c0[:, 0])),
           "Q": d[:, 2] / (np.abs(c0[:, 2]) + c0[:, 1] ** 2)}
    out = {}
    for k, v in rel.items():
        out[k] = {"median": float(np.median(v)), "p99": float(np.quantile(v, 0.99))}
    return out


def _host_at(torch, outs, idx=None):
    """Per-ray outputs copied to the host (all rays, or the rays of idx)."""
    if idx is None:
        h = {k: v.cpu().numpy() for k, v in outs.items()}
    else:
        it = torch.from_numpy(np.asarray(idx, dtype=np.int64)).to(next(iter(outs.values())).device)
        h = {k: v.index_select(0, it).cpu().numpy() for k, v in outs.items()}
    h["attempts"] = h["attempts"].astype(np.uint32)
    return h


def _roofline_frac(stats, kernel_ms):
    f32, f64 = flops(stats, "kerr")
    peak = (f32 + f64) / (f32 / PEAK_FP32 + f64 / PEAK_FP64) if (f32 + f64) else PEAK_FP64
    return (f32 + f64) / (kernel_ms * 1e-3) / 1e12 / peak, (f32 + f64) / (kernel_ms * 1e-3) / 1e12


def _vs_oracle(gpu, ora, band=None):
    """Class agreement (outside the band too when given), final states within 1e-6, attempt counts, PSNR of
    the disk intensity over the sampled rays."""
    from tests.parity import psnr, state_rel_err
    eq = gpu["cls"] == ora["cls"]
    err = state_rel_err(gpu, ora)
    out = {"rays": int(len(eq)), "class_agree": float(np.mean(eq)),
           "err_le_1e-6": float(np.mean(err[eq] <= 1e-6)) if eq.any() else 1.0,
           "err_median": float(np.median(err[eq])) if eq.any() else 0.0,
           "attempts_equal": float(np.mean(gpu["attempts"] == ora["attempts"])),
           "psnr_intensity": psnr(gpu["intensity"], ora["intensity"])}
    if not math.isfinite(out["psnr_intensity"]):
        out["psnr_intensity"] = None      # identical intensities
    if band is not None:
        ok = eq & ~band
        out["class_agree_outside_band"] = float(np.mean(eq[~band])) if (~band).any() else 1.0
        out["err_le_1e-6_outside_band"] = float(np.mean(err[ok] <= 1e-6)) if ok.any() else 1.0
    return out


def _stride_idx(w, h, stride):
    return np.array([py * w + px for py in range(0, h, stride) for px in range(0, w, stride)], dtype=np.int64)


def frame_kerr4096(torch, rend, size=4096, ora_stride=32):
    """One timed frame of the near-extremal workload per precision (a first frame in tile order, then the
    frame under the learned longest-first order), with roofline fractions and the oracle on a stride sample."""
    import gargantua_amd as gg
    from oracle import Oracle
    from tests.parity import boundary_band
    scene = gg.configs.kerr_4096(size, size)
    n = scene.n_rays
    idx = _stride_idx(size, size, ora_stride)
    res, sub = {}, {}
    band = None
    for prec in ("fp64", "mixed"):
        outs = _dev_outputs(torch, n)
        rend.set_launch(64, 0, 64, 1, 1)
        first = rend.render(scene, prec, outputs=dict(_ptrs(outs)))["stats"]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = rend.render(scene, prec, outputs=dict(_ptrs(outs)))["stats"]
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        frac, tf = _roofline_frac(st, st["kernel_ms"])
        res[prec] = {"value": n / wall, "kernel_ms": st["kernel_ms"], "first_frame_kernel_ms": first["kernel_ms"],
                     "roofline_frac": frac, "achieved_tflops": tf, "attempts_per_ray": st["attempts"] / n,
                     "fp32_rhs_share": st["rhs_evals_fp32"] / st["rhs_evals"],
                     "fast_rhs_share": st["rhs_evals_fast"] / st["rhs_evals"]}
        if prec == "fp64":
            band = boundary_band(outs["cls"].cpu().numpy().reshape(size, size)).ravel()[idx]
        sub[prec] = _host_at(torch, outs, idx)
        _log(f"kerr4096 {prec}: first frame {first['kernel_ms']:.0f} ms, learned order {st['kernel_ms']:.0f} ms")
        del outs
        torch.cuda.empty_cache()
    rend.set_launch()
    ora = Oracle(threads=max(1, min(16, len(os.sched_getaffinity(0)))))
    t0 = time.perf_counter()
    ref = ora.render_rays(scene, idx.astype(np.uint64), state=False)
    dt = time.perf_counter() - t0
    return {"workload": f"Kerr a=0.998 {size}x{size}, camera r=1000M i=85deg FOV 20deg, escape 1100M, 50000 attempts",
            "launch": "64 threads x 12 blocks per CU (the default for this many rays), refill 64, learned order", "rays": n,
            "fp64": res["fp64"], "mixed": res["mixed"],
            "mixed_speedup": res["fp64"]["kernel_ms"] / res["mixed"]["kernel_ms"],
            "fp64_vs_oracle": _vs_oracle(sub["fp64"], ref, band), "mixed_vs_oracle": _vs_oracle(sub["mixed"], ref, band),
            "oracle_sample": f"every {ora_stride}th row and column ({len(idx)} rays, {dt:.1f} s on {ora.threads} threads)"}


def frame_batch64(torch, rend, ora_stride=16):
    """One timed frame of the 64-scene batch per precision (graph-captured launch, learned order after a first
    frame), with roofline fractions and the oracle on a stride sample of every scene."""
    import gargantua_amd as gg
    from oracle import Oracle
    from tests.parity import boundary_band
    scenes = gg.configs.batch_64(seed=2026)
    n = sum(s.n_rays for s in scenes)
    idxs = [_stride_idx(s.width, s.height, ora_stride) for s in scenes]
    res, sub, bands = {}, {}, []
    for prec in ("fp64", "mixed"):
        outs = [_dev_outputs(torch, s.n_rays) for s in scenes]
        rend.set_launch(64, 0, 64, 1, 1)
        _, first = rend.render_batch(scenes, prec, outputs=[dict(_ptrs(o)) for o in outs])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, st = rend.render_batch(scenes, prec, outputs=[dict(_ptrs(o)) for o in outs])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        frac, tf = _roofline_frac(st, st["kernel_ms"])
        res[prec] = {"value": n / wall, "kernel_ms": st["kernel_ms"], "first_frame_kernel_ms": first["kernel_ms"],
                     "roofline_frac": frac, "achieved_tflops": tf, "attempts_per_ray": st["attempts"] / n,
                     "fp32_rhs_share": st["rhs_evals_fp32"] / st["rhs_evals"],
                     "fast_rhs_share": st["rhs_evals_fast"] / st["rhs_evals"]}
        if prec == "fp64":
            bands = [boundary_band(o["cls"].cpu().numpy().reshape(s.height, s.width)).ravel()[ix]
                     for o, s, ix in zip(outs, scenes, idxs)]
        sub[prec] = [_host_at(torch, o, ix) for o, ix in zip(outs, idxs)]
        _log(f"batch64 {prec}: first frame {first['kernel_ms']:.0f} ms, learned order {st['kernel_ms']:.0f} ms")
        del outs
        torch.cuda.empty_cache()
    rend.set_launch()
    ora = Oracle(threads=max(1, min(16, len(os.sched_getaffinity(0)))))
    t0 = time.perf_counter()
    refs = [ora.render_rays(s, ix.astype(np.uint64), state=False) for s, ix in zip(scenes, idxs)]
    dt = time.perf_counter() - t0
    cat = lambda parts: {k: np.concatenate([p[k] for p in parts]) for k in ("cls", "r", "theta", "phi", "attempts", "intensity")}
    ref_all, band_all = cat(refs), np.concatenate(bands)
    worst = {p: min(_vs_oracle(g, r)["err_le_1e-6"] for g, r in zip(sub[p], refs)) for p in sub}
    return {"workload": "64 seeded Kerr scenes (generator seed 2026) in one graph-captured launch",
            "launch": "64 threads x 12 blocks per CU (the default for this many rays), refill 64, learned order", "rays": n,
            "fp64": res["fp64"], "mixed": res["mixed"],
            "mixed_speedup": res["fp64"]["kernel_ms"] / res["mixed"]["kernel_ms"],
            "fp64_vs_oracle": {**_vs_oracle(cat(sub["fp64"]), ref_all, band_all),
                               "worst_scene_err_le_1e-6": worst["fp64"]},
            "mixed_vs_oracle": {**_vs_oracle(cat(sub["mixed"]), ref_all, band_all),
                                "worst_scene_err_le_1e-6": worst["mixed"]},
            "oracle_sample": f"every {ora_stride}th row and column of each scene ({len(band_all)} rays, {dt:.1f} s "
                             f"on {ora.threads} threads)"}


def block_adjoint(torch, rend, steps=5, warmup=2, fd_size=40):
    """Adjoint gradients/s on the 512^2 differentiable render (mixed FP32/FP64 forward recording every attempt,
    learned ray order, target resident on the device), and the gradient's error against central differences of
    the oracle's loss on the same scene at fd_size^2 (h = 1e-6 over every ray; h = 1e-5 over the rays whose step
    decisions do not flip)."""
    import gargantua_amd as gg
    from oracle import Oracle
    from tests.gradcheck import central_fd, masked_gradient
    cfg = gg.configs
    tgt = cfg.grad_512_params(512, 512, spin=0.7, inclination_deg=60.0, q=2.5, r_in=cfg.r_isco(1.0, 0.7))
    target = torch.from_numpy(rend.grad_image(tgt)).to("cuda")
    gm = cfg.grad_512_params(512, 512, precision="mixed")
    rend.set_launch(order=1)
    for _ in range(warmup):
        rend.grad(gm, target, checkpoint_every=0, want_image=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sts = []
    for _ in range(steps):
        out = rend.grad(gm, target, checkpoint_every=0, want_image=False)
        sts.append(out["stats"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    rend.set_launch()
    st = sts[-1]
    # finite differences of the oracle's loss on the small scene
    ora = Oracle(threads=max(1, min(16, len(os.sched_getaffinity(0)))))
    mk = lambda: cfg.grad_512_params(fd_size, fd_size)
    t_small = ora.grad(cfg.grad_512_params(fd_size, fd_size, spin=0.7, inclination_deg=60.0, q=2.5,
                                           r_in=cfg.r_isco(1.0, 0.7)), np.zeros(fd_size * fd_size))["intensity"]
    fd6 = central_fd(ora, mk, t_small, 1e-6)
    fd5 = central_fd(ora, mk, t_small, 1e-5)
    fd_err = {}
    for prec in ("fp64", "mixed"):
        a = rend.grad(cfg.grad_512_params(fd_size, fd_size, precision=prec), t_small, want_dI=True)
        ad5 = masked_gradient(a["dI"], a["intensity"], t_small, fd5["mask"])
        fd_err[prec] = {"h1e-6_max_rel": float(np.max(np.abs(a["grad"] - fd6["fd"]) / np.abs(fd6["fd"]))),
                        "h1e-5_smooth_rays_max_rel": float(np.max(np.abs(ad5 - fd5["fd_smooth"]) /
                                                                  np.abs(fd5["fd_smooth"]))),
                        "grad": a["grad"].tolist()}
    return {"workload": "512x512 Kerr thin disk r_out 20M edge 0.25M, camera r=50M FOV 30deg, rtol 1e-8 atol 1e-10; "
                        "theta0 (0.5, 45deg, 3.0, r_ISCO+0.5), target (0.7, 60deg, 2.5, r_ISCO(0.7))",
            "grads_per_s": steps / wall, "ms_per_grad": 1e3 * wall / steps,
            "forward_ms": float(np.mean([x["forward_ms"] for x in sts])),
            "adjoint_ms": float(np.mean([x["adjoint_ms"] for x in sts])),
            "mode": st["mode"], "checkpoint_every": st["checkpoint_every"], "forward_precision": "mixed",
            "disk_rays": st["disk_rays"], "attempts_backpropagated": st["steps"],
            "step_sizes": "differentiated through the error estimate and controller",
            "fd_check": {"scene": f"the same scene at {fd_size}x{fd_size}", "fd_h1e-6": fd6["fd"].tolist(),
                         "fd_h1e-5_smooth_rays": fd5["fd_smooth"].tolist(),
                         "rays_with_decision_flips_h1e-5": fd5["flips"].tolist(), **fd_err}}


def bench_batch64(args):
    """64 seeded scenes (a~U[0,0.998], i~U[5,89] deg, r_cam~U[30,500], res in {128..1024}) in one
    hipGraph-captured persistent launch after the autotuner picked the launch shape."""
    import torch
    import gargantua_amd as gg
    scenes = gg.configs.batch_64(seed=2026)
    n = sum(s.n_rays for s in scenes)
    rend = gg.Renderer(0, stream=torch.cuda.current_stream().cuda_stream)
    outs = [_ptrs(_dev_outputs(torch, s.n_rays)) for s in scenes]
    # tune each precision on every 8th scene (same distribution, an eighth of the rays)
    tuned, tune_s = {}, {}
    res = {}
    for prec in ("mixed", "fp64"):
        t0 = time.perf_counter()
        tuned[prec] = rend.autotune(scenes[::8], prec)
        tune_s[prec] = time.perf_counter() - t0
        _log(f"autotune {prec}: {tuned[prec][0]} ({tuned[prec][1]:.1f} ms on {len(scenes[::8])} scenes) "
             f"in {tune_s[prec]:.1f} s")
        L = tuned[prec][0]
        rend.set_launch(L["block_threads"], L["blocks_per_cu"], L["refill_interval"], 1, L["order"])
        for _ in range(args.warmup):
            rend.render_batch(scenes, prec, outputs=[dict(o) for o in outs])
        torch.cuda.synchronize()
        kms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            _, st = rend.render_batch(scenes, prec, outputs=[dict(o) for o in outs])
            kms.append(st["kernel_ms"])
            _log(f"batch64 {prec}: {st['kernel_ms']:.1f} ms")
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        res[prec] = {"value": n * args.steps / wall, "kernel_ms": float(np.median(kms)),
                     "attempts_per_ray": st["attempts"] / n, "fp32_rhs_share": st["rhs_evals_fp32"] / st["rhs_evals"]}
    rend.close()
    return {"metric": "geodesics/s, 64 seeded Kerr scenes in one graph-captured launch", "value": res["mixed"]["value"],
            "unit": "geodesics/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 * n / res["mixed"]["value"], "higher_is_better": True, "dtype": "f32+f64",
            "config": {"workload": "batch of 64 scenes, generator seed 2026", "rays": n,
                       "resolutions": [s.width for s in scenes],
                       "autotuned_launch": {p: tuned[p][0] for p in tuned},
                       "autotune_best_ms": {p: tuned[p][1] for p in tuned}, "autotune_seconds": tune_s},
            "mixed": res["mixed"], "fp64": res["fp64"], "mixed_speedup": res["fp64"]["kernel_ms"] / res["mixed"]["kernel_ms"]}


def bench_kerr4096(args):
    """Near-extremal Kerr a=0.998 at 4096x4096, camera r=1000M, i=85 deg, FOV 20 deg: mixed and FP64,
    classification agreement inside and outside a +-2 px band around class boundaries."""
    import torch
    import gargantua_amd as gg
    from tests.parity import compare
    size = args.size if args.size != 1024 else 4096
    scene = gg.configs.kerr_4096(size, size)
    n = scene.n_rays
    rend = gg.Renderer(0, stream=torch.cuda.current_stream().cuda_stream)
    res, host = {}, {}
    for prec in ("mixed", "fp64"):
        outs = _dev_outputs(torch, n)
        wall, kms, st = time_precision(torch, rend, scene, prec, 1, 0, outs)
        _log(f"kerr4096 {prec}: {kms[0]:.1f} ms")
        res[prec] = {"value": n / wall, "kernel_ms": kms[0], "attempts_per_ray": st["attempts"] / n,
                     "max_attempts": None, "fp32_rhs_share": st["rhs_evals_fp32"] / st["rhs_evals"]}
        h = {k: v.cpu().numpy() for k, v in outs.items()}
        h["attempts"] = h["attempts"].astype(np.uint32)
        res[prec]["max_attempts"] = int(h["attempts"].max())
        res[prec]["classes"] = np.bincount(h["cls"], minlength=4).tolist()
        host[prec] = h
        del outs
        torch.cuda.empty_cache()
    cmp = compare(host["mixed"], host["fp64"], shape=(size, size))
    rend.close()
    return {"metric": f"geodesics/s Kerr a=0.998 {size}²", "value": res["mixed"]["value"], "unit": "geodesics/s",
            "n_gpus": 1, "dtype": "f32+f64", "config": {"workload": f"Kerr a=0.998 {size}x{size}, camera r=1000M "
                                                        "i=85deg FOV 20deg, escape 1100M, 50000 attempts"},
            "mixed": res["mixed"], "fp64": res["fp64"], "mixed_speedup": res["fp64"]["kernel_ms"] / res["mixed"]["kernel_ms"],
            "mixed_vs_fp64": cmp}


def bench_scene(args):
    """The reference CLI's lensed-spheres demo (main.rs:78-202): Schwarzschild M=1, camera r=30M looking at the
    hole, four Lambertian spheres and a light behind it, 800x600, without and with the volumetric disk."""
    import torch
    import gargantua_amd as gg
    from oracle import Oracle
    w, h = (800, 600) if args.size == 1024 else (args.size, args.size * 3 // 4)
    world = gg.configs.lensed_spheres_world()
    cam = gg.Camera(position=(0.0, 30.0, math.pi / 2, 0.0), look_at=(-1.0, 0.0, 0.0), up=(0.0, 0.0, 1.0),
                    fov_y_radians=math.radians(60.0), aspect=w / h)
    rend = gg.Renderer(0, stream=torch.cuda.current_stream().cuda_stream)
    rend.set_hittables(world)
    out = {}
    for name, disk in (("scene", None), ("disk_scene", gg.AccretionDisk())):
        scene = gg.Scene(metric=gg.Schwarzschild(1.0), camera=cam, disk=disk,
                         integrator=gg.RK45Integrator(max_radius=200.0, max_steps=5000), width=w, height=h)
        n = scene.n_rays
        res, host = {}, {}
        for prec in ("mixed", "fp64"):
            outs = _dev_outputs(torch, n)
            outs["color"] = torch.empty(3 * n, dtype=torch.float64, device="cuda")
            wall, kms, st = time_precision(torch, rend, scene, prec, max(1, args.steps), 1, outs)
            kms = sorted(kms)
            res[prec] = {"geodesics_per_s": n / (wall / max(1, args.steps)), "kernel_ms": kms[len(kms) // 2],
                         "attempts_per_ray": st["attempts"] / n}
            host[prec] = {k: v.cpu().numpy() for k, v in outs.items()}
            _log(f"scene {name} {prec}: {res[prec]['kernel_ms']:.2f} ms")
        cm, cf = host["mixed"]["cls"], host["fp64"]["cls"]
        hit = (cf == gg.GG_CLASS_SCENE) & (cm == cf)
        col = host["mixed"]["color"].reshape(-1, 3)[hit] - host["fp64"]["color"].reshape(-1, 3)[hit]
        res["classes"] = np.bincount(cf, minlength=5).tolist()
        res["mixed_vs_fp64"] = {"class_agree": float(np.mean(cm == cf)),
                                "scene_color_max_abs_diff": float(np.abs(col).max()) if hit.any() else 0.0}
        res["mixed_speedup"] = res["fp64"]["kernel_ms"] / res["mixed"]["kernel_ms"]
        # the same rays without the scene: what the chord tests cost on top
        rend.set_hittables(None)
        for prec in ("mixed", "fp64"):
            outs = _dev_outputs(torch, n)
            _, kms, _ = time_precision(torch, rend, scene, prec, max(1, args.steps), 1, outs)
            res[prec]["kernel_ms_without_scene"] = sorted(kms)[len(kms) // 2]
        rend.set_hittables(world)
        if not args.no_cpu:
            ora = Oracle(threads=max(1, min(16, len(os.sched_getaffinity(0)))))
            ora.set_hittables(world)
            idx = np.array([py * w + px for py in range(0, h, 8) for px in range(0, w, 8)], dtype=np.uint64)
            t0 = time.perf_counter()
            ora.render_rays(scene, idx, state=False)
            dt = time.perf_counter() - t0
            ora.set_hittables(None)
            res["cpu_baseline"] = {"value": len(idx) / dt, "unit": "geodesics/s", "cores": ora.threads,
                                   "kind": "port", "sample": f"every 8th row and column ({len(idx)} rays)"}
        out[name] = res
    # BVH scaling: 4096 small spheres scattered in a shell 25-90M around the hole (seeded)
    rng = np.random.default_rng(2026)
    many = gg.World()
    mats = [gg.Lambertian((0.8, 0.3, 0.2)), gg.Lambertian((0.2, 0.5, 0.9)), gg.DiffuseLight((1.5, 1.3, 0.9))]
    for i in range(4096):
        d = rng.normal(size=3)
        c = d / np.linalg.norm(d) * rng.uniform(25.0, 90.0)
        if c[0] > 20.0 and abs(c[1]) < 15.0 and abs(c[2]) < 15.0:
            c[0] = -c[0]                                   # keep the camera's neighbourhood clear
        many.add(gg.Sphere(tuple(c), float(rng.uniform(0.3, 1.0)), mats[i % 3]))
    rend.set_hittables(many)
    scene = gg.Scene(metric=gg.Schwarzschild(1.0), camera=cam, integrator=gg.RK45Integrator(max_radius=200.0,
                                                                                       max_steps=5000),
                     width=w, height=h)
    n = scene.n_rays
    res = {"primitives": len(many)}
    for prec in ("mixed", "fp64"):
        outs = _dev_outputs(torch, n)
        outs["color"] = torch.empty(3 * n, dtype=torch.float64, device="cuda")
        _, kms, st = time_precision(torch, rend, scene, prec, max(1, args.steps), 1, outs)
        res[prec] = {"kernel_ms": sorted(kms)[len(kms) // 2], "attempts_per_ray": st["attempts"] / n}
        res["classes"] = np.bincount(outs["cls"].cpu().numpy(), minlength=5).tolist()
        _log(f"scene many_spheres {prec}: {res[prec]['kernel_ms']:.2f} ms")
    out["many_spheres"] = res
    rend.set_hittables(None)
    rend.close()
    return {"metric": f"geodesics/s lensed-spheres {w}x{h}", "value": out["scene"]["mixed"]["geodesics_per_s"],
            "unit": "geodesics/s", "n_gpus": 1, "dtype": "f32+f64",
            "config": {"workload": f"lensed-spheres {w}x{h} Schwarzschild M=1, camera r=30M FOV 60deg, "
                                   "5 spheres, max_radius 200M, 5000 attempts"},
            **out}


def bench_user(args):
    """Metrics known only by their tensor, compiled at run time: Reissner-Nordstrom Q=0.6 and Kerr a=0.9 written as
    tensors, 512x512 thin disk, camera r=30M i=75deg FOV 40deg, FP64; the Kerr tensor beside the library's
    closed-form Kerr kernel on the same scene."""
    import torch
    import gargantua_amd as gg
    from tests.test_user_metric import KERR, SYMMETRIC, rn_metric
    size = 512 if args.size == 1024 else args.size
    cam = gg.configs.inclination_camera(30.0, math.radians(75.0), math.radians(40.0))
    disk = gg.AccretionDisk(kind=gg.GG_DISK_THIN, r_inner=gg.configs.r_isco(1.0, 0.9), r_outer=20.0)
    rend = gg.Renderer(0, stream=torch.cuda.current_stream().cuda_stream)
    out = {}
    rh_kerr = 1.0 + math.sqrt(1.0 - 0.81)
    rn = rn_metric(0.6)
    metrics = {"reissner_nordstrom": rn,
               "reissner_nordstrom_symmetric": gg.UserMetric(SYMMETRIC + rn.source, rn.mass, rn.spin, rn.event_horizon),
               "kerr_tensor": gg.UserMetric(KERR, mass=1.0, spin=0.9, event_horizon=rh_kerr),
               "kerr_tensor_symmetric": gg.UserMetric(SYMMETRIC + KERR, mass=1.0, spin=0.9, event_horizon=rh_kerr),
               "kerr_closed_form": gg.Kerr(1.0, 0.9)}
    for name, m in metrics.items():
        scene = gg.Scene(metric=m, camera=cam, disk=disk, integrator=gg.RK45Integrator(max_radius=200.0),
                         width=size, height=size)
        t0 = time.perf_counter()
        rend._ensure_metric(scene)
        compile_s = time.perf_counter() - t0
        outs = _dev_outputs(torch, scene.n_rays)
        wall, kms, st = time_precision(torch, rend, scene, "fp64", max(1, args.steps), 1, outs)
        kms = sorted(kms)
        out[name] = {"kernel_ms": kms[len(kms) // 2], "geodesics_per_s": scene.n_rays / (wall / max(1, args.steps)),
                     "attempts_per_ray": st["attempts"] / scene.n_rays, "rhs_evals": st["rhs_evals"],
                     "first_set_s": compile_s}
        _log(f"user {name}: {out[name]['kernel_ms']:.1f} ms")
    rend.close()
    if not args.no_cpu:
        # the FP64 oracle (the same tensor compiled for the host) on every 8th row and column
        from oracle import Oracle
        ora = Oracle(threads=max(1, min(16, len(os.sched_getaffinity(0)))))
        ora.set_user_metric(rn)
        scene = gg.Scene(metric=rn, camera=cam, disk=disk, integrator=gg.RK45Integrator(max_radius=200.0),
                         width=size, height=size)
        idx = np.array([py * size + px for py in range(0, size, 8) for px in range(0, size, 8)], dtype=np.uint64)
        t0 = time.perf_counter()
        ora.render_rays(scene, idx, state=False)
        dt = time.perf_counter() - t0
        ora.set_user_metric(None)
        out["cpu_baseline"] = {"value": len(idx) / dt, "unit": "geodesics/s", "cores": ora.threads, "kind": "port",
                               "sample": f"Reissner-Nordstrom tensor, every 8th row and column ({len(idx)} rays)"}
    out["tensor_vs_closed_form"] = out["kerr_tensor"]["kernel_ms"] / out["kerr_closed_form"]["kernel_ms"]
    out["symmetric_tensor_vs_closed_form"] = (out["kerr_tensor_symmetric"]["kernel_ms"] /
                                              out["kerr_closed_form"]["kernel_ms"])
    return {"metric": f"geodesics/s user metric {size}x{size}", "value": out["reissner_nordstrom"]["geodesics_per_s"],
            "unit": "geodesics/s", "n_gpus": 1, "dtype": "f64",
            "config": {"workload": f"{size}x{size} thin disk, camera r=30M i=75deg FOV 40deg; metrics given as tensors "
                                   "(numerical Christoffels, h=1e-6)"}, **out}


# Algorithmic FLOPs of one accepted step back-propagated by the adjoint
# reverse kernel: six reverse-mode RHS vector-Jacobian products (628 each,
# counted from the source of kerr_rhs_vjp: 130 for the forward symbol
# algebra, 342 for the backward sweep, 52 for the contraction coefficients,
# 104 for the velocity gradient) and ~250 for the stage cotangents and their
# combinations.  The stage velocities come from the step record, so no RHS
# is evaluated.  Re-traces and segment recomputations are counted like the
# trace kernel (237 per RHS, 318 per full attempt).
F_VJP = 628
F_ADJ_STEP = 6 * F_VJP + 250


def bench_grad512(args):
    """Differentiable render: 512x512 Kerr thin disk at theta0 = (a=0.5, i=45 deg, q=3, r_in=r_ISCO(0.5)+0.5),
    loss = MSE against the image at (0.7, 60 deg, 2.5, r_ISCO(0.7)); dL/dtheta by the adjoint kernels with
    the render recording every step (K = 0, automatic; FP64, and mixed FP32/FP64 as the headline), with the
    disk rays re-traced and every step stored (K = 65536), and with checkpoints every K in {8, 32, 128}
    accepted steps; then gradient-descent (Adam, mixed forward) recovery of theta."""
    import torch
    import gargantua_amd as gg
    cfg = gg.configs
    size = args.size if args.size != 1024 else 512
    rend = gg.Renderer(0, stream=torch.cuda.current_stream().cuda_stream)
    tgt_scene = cfg.grad_512_params(size, size, spin=0.7, inclination_deg=60.0, q=2.5, r_in=cfg.r_isco(1.0, 0.7))
    target = rend.grad_image(tgt_scene)
    target_dev = torch.from_numpy(target).to("cuda")   # resident in HBM, read in place
    g0 = cfg.grad_512_params(size, size)
    per_k = {}
    g0m = cfg.grad_512_params(size, size, precision="mixed")
    # (label, K, learned ray order for the render, forward precision)
    # the headline run last, on a warm GPU
    runs = [("auto", 0, 0, g0), ("K65536", 65536, 0, g0), ("K8", 8, 0, g0), ("K32", 32, 0, g0), ("K128", 128, 0, g0),
            ("auto+order", 0, 1, g0), ("auto+order mixed", 0, 1, g0m)]
    for K, (label, kk, order, gs) in [(r[0], r) for r in runs]:
        rend.set_launch(order=order)
        for _ in range(args.warmup):
            rend.grad(gs, target_dev, checkpoint_every=kk, want_image=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sts = []
        for _ in range(args.steps):
            out = rend.grad(gs, target_dev, checkpoint_every=kk, want_image=False)
            sts.append(out["stats"])
        wall = time.perf_counter() - t0
        st = sts[-1]
        adj_ms = float(np.mean([x["adjoint_ms"] for x in sts]))
        fl = st["steps"] * F_ADJ_STEP + st["attempts"] * (6 * 237 + 318)
        per_k[K] = {"grads_per_s": args.steps / wall, "ms_per_grad": 1e3 * wall / args.steps,
                    "forward_ms": float(np.mean([x["forward_ms"] for x in sts])), "adjoint_ms": adj_ms,
                    "adjoint_tflops": fl / (adj_ms * 1e-3) / 1e12, "adjoint_frac_fp64_peak":
                        fl / (adj_ms * 1e-3) / 1e12 / PEAK_FP64,
                    "disk_rays": st["disk_rays"], "steps_backpropagated": st["steps"],
                    "attempts_retraced": st["attempts"], "checkpoints": st["checkpoints"],
                    "checkpoint_MiB": st["checkpoint_bytes"] / 2**20, "segment_MiB": st["segment_bytes"] / 2**20,
                    "record_MiB": st["record_bytes"] / 2**20, "checkpoint_every": st["checkpoint_every"],
                    "mode": st["mode"], "learned_order": bool(order), "total_ms_in_library":
                        float(np.mean([x["total_ms"] for x in sts])),
                    "grid_blocks": st["grid_blocks"], "loss": out["loss"], "grad": out["grad"].tolist()}
        _log(f"grad512 K={K}: {per_k[K]['ms_per_grad']:.1f} ms/grad (adjoint {adj_ms:.1f} ms, "
             f"forward {per_k[K]['forward_ms']:.1f} ms, mode {st['mode']})")
    rend.set_launch(order=1)
    same = all(per_k[K]["grad"] == per_k["K8"]["grad"] for K in per_k if "mixed" not in K)
    mixed_rel = (np.abs(np.array(per_k["auto+order mixed"]["grad"]) - np.array(per_k["K8"]["grad"])) /
                 np.abs(np.array(per_k["K8"]["grad"]))).max()
    # gradient-descent recovery (Adam on the raw parameters)
    gd = None
    if args.gd_steps > 0:
        t0 = time.perf_counter()
        fit = rend.fit(g0m, target_dev, args.gd_steps,
                       log=lambda it, loss, th: _log(f"gd {it}: loss {loss:.3e} theta {np.round(th, 4).tolist()}"))
        theta = fit["theta"][-1]
        truth = np.array([tgt_scene.spin, tgt_scene.inclination, tgt_scene.emissivity_q, tgt_scene.r_inner])
        gd = {"steps": args.gd_steps, "optimizer": "Adam lr 0.02 (0.9, 0.999)", "seconds": time.perf_counter() - t0,
              "theta0": [g0.spin, g0.inclination, g0.emissivity_q, g0.r_inner], "theta_final": theta.tolist(),
              "theta_true": truth.tolist(), "abs_err": np.abs(theta - truth).tolist(),
              "loss_first": float(fit["loss"][0]), "loss_last": float(fit["loss"][-1])}
    rend.close()
    best = per_k["auto+order mixed"]
    return {"metric": f"adjoint gradient evaluations/s, {size}x{size} Kerr thin disk, dL/d(a, i, q, r_in)",
            "value": best["grads_per_s"], "unit": "gradients/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": best["ms_per_grad"], "higher_is_better": True, "dtype": "f64",
            "data": "synthetic (target image rendered at the true parameters)",
            "config": {"workload": f"{size}x{size} Kerr thin disk r_out 20M edge 0.25M, camera r=50M FOV 30deg, "
                                   "rtol 1e-8 atol 1e-10; theta0 (0.5, 45deg, 3.0, r_ISCO+0.5), target "
                                   "(0.7, 60deg, 2.5, r_ISCO(0.7)); value: mixed FP32/FP64 forward render "
                                   "recording every step (K=0), learned ray order, target resident on the device",
                       "K": 0, "forward_precision": "mixed"},
            "roofline": {"bound": "valu", "achieved": best["adjoint_tflops"], "peak": PEAK_FP64, "unit": "TFLOP/s",
                         "frac": best["adjoint_frac_fp64_peak"],
                         "traffic": profiled_traffic("adjoint_grad512_counters.md", "gg::adjoint_rev_chunked_kernel"),
                         "traffic_source": "profiles/adjoint_grad512_counters.md, adjoint_rev_chunked_kernel "
                                           "(dominant kernel; reads 4.4 GB of step records)",
                         "kernel": "adjoint_seed_kernel + adjoint_rev_chunked_kernel + adjoint_cam_kernel"},
            "per_K": per_k, "grad_identical_across_K": same, "mixed_forward_grad_max_rel_diff": float(mixed_rel),
            "gradient_descent": gd}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--workload", default="kerr1024", choices=["kerr1024", "batch64", "kerr4096", "grad512", "scene", "user", "copy"],
                    help="copy: one 512 MiB device-to-device copy, the reference kernel that calibrates the "
                         "FETCH_SIZE/WRITE_SIZE counters for HBM traffic")
    ap.add_argument("--gd-steps", type=int, default=200, help="grad512: gradient-descent iterations (0: none)")
    ap.add_argument("--headline-only", action="store_true",
                    help="kerr1024: only the 1024^2 workload (no 4096^2 / batch frames, no adjoint block)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              "(rays are independent; the multi-GPU split is described in the design notes)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch64":
        print(json.dumps(bench_batch64(args)))
        return
    if args.workload == "kerr4096":
        print(json.dumps(bench_kerr4096(args)))
        return
    if args.workload == "grad512":
        print(json.dumps(bench_grad512(args)))
        return
    if args.workload == "scene":
        print(json.dumps(bench_scene(args)))
        return
    if args.workload == "user":
        print(json.dumps(bench_user(args)))
        return
    if args.workload == "copy":
        import torch
        x = torch.ones(64 * 2**20, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"workload": "copy", "bytes_read": x.numel() * 8, "bytes_written": x.numel() * 8}))
        return

    import torch
    import gargantua_amd as gg
    from tests.parity import compare

    scene = gg.configs.kerr_1024(args.size, args.size)
    n = scene.n_rays
    rend = gg.Renderer(0, stream=torch.cuda.current_stream().cuda_stream)
    outs_m = _dev_outputs(torch, n)
    outs_d = _dev_outputs(torch, n)

    # Launch shape per precision from the autotuner: block size, blocks per CU,
    # and whether rays are queued longest first by the previous render's
    # attempt counts (every ray is traced in full either way).
    cands = [(64, 12, 64, 2, 0), (64, 12, 64, 2, 1), (64, 12, 16, 2, 1), (64, 8, 64, 2, 1)]
    tuned = {}
    for prec in ("mixed", "fp64", "fp64fast"):
        tuned[prec] = rend.autotune([scene], prec, candidates=cands)[0]
        _log(f"kerr1024 {prec}: tuned launch {tuned[prec]}")
    tile_ms = {}

    def run(prec, outs):
        L = tuned[prec]
        rend.set_launch(L["block_threads"], L["blocks_per_cu"], L["refill_interval"], 1, 0)
        tile_ms[prec] = rend.render(scene, prec, outputs=dict(_ptrs(outs)))["stats"]["kernel_ms"]   # first frame
        rend.set_launch(L["block_threads"], L["blocks_per_cu"], L["refill_interval"], 1, L["order"])
        return time_precision(torch, rend, scene, prec, args.steps, args.warmup, outs)

    wall_m, kms_m, st_m = run("mixed", outs_m)
    wall_d, kms_d, st_d = run("fp64", outs_d)
    outs_f = _dev_outputs(torch, n)
    wall_f, kms_f, st_f = run("fp64fast", outs_f)

    # accuracy of the mixed kernel against the FP64 kernel (same inputs)
    host = lambda o: {k: v.cpu().numpy() for k, v in o.items()}
    hm, hd, hf = host(outs_m), host(outs_d), host(outs_f)
    for h in (hm, hd, hf):
        h["attempts"] = h["attempts"].astype(np.uint32)
    cmp = compare(hm, hd, shape=(scene.height, scene.width))
    cmp_f = compare(hf, hd, shape=(scene.height, scene.width))

    value = n * args.steps / wall_m
    ms_step = 1e3 * wall_m / args.steps
    k_ms = float(np.median(kms_m))   # median of the timed launches
    f32, f64 = flops(st_m, "kerr")
    tflops = (f32 + f64) / (k_ms * 1e-3) / 1e12
    peak = (f32 + f64) / (f32 / PEAK_FP32 + f64 / PEAK_FP64) if (f32 + f64) else PEAK_FP64
    d32, d64 = flops(st_d, "kerr")
    k_ms_d = float(np.median(kms_d))

    line = {
        "metric": METRIC, "value": value, "unit": "geodesics/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic (camera rays of the scene; no inputs)",
        "config": {"workload": f"Kerr a=0.9 {args.size}x{args.size} thin disk r_ISCO..30M q=3, camera r=100M "
                               "i=60deg FOV 25deg, jitter seed 2026, rtol 1e-8 atol 1e-10, escape 1000M, "
                               "capture 1.01 r_h, 20000 attempts", "precision": "mixed",
                   "rays": n, "launch": tuned["mixed"],
                   "schedule": "autotuned launch; order 1 = rays queued longest first by the previous render's "
                               "attempt counts (learned in the warm-up), every ray traced in full",
                   "first_frame_tile_order_ms": tile_ms},
        "roofline": {"bound": "valu", "achieved": tflops, "peak": peak, "unit": "TFLOP/s", "frac": tflops / peak,
                     "traffic": profiled_traffic("trace_kernel_kerr1024_counters.md", "gg::trace_kernel<1, 2, 1, 0, 0>"),
                     "traffic_source": "profiles/trace_kernel_kerr1024_counters.md (rocprofv3 FETCH_SIZE and "
                                       "WRITE_SIZE, copy-calibrated); algorithmic outputs 47 MB",
                     "kernel": "trace_kernel<Kerr, thin, mixed>",
                     "kernel_ms": k_ms, "flop_per_launch": f32 + f64, "fp32_flop_share": f32 / (f32 + f64)},
        "fp64": {"value": n * args.steps / wall_d, "kernel_ms": k_ms_d,
                 "roofline_frac": (d32 + d64) / (k_ms_d * 1e-3) / 1e12 / PEAK_FP64,
                 "mixed_speedup": k_ms_d / k_ms},
        "fp64fast": {"value": n * args.steps / wall_f, "kernel_ms": float(np.median(kms_f)),
                     "vs_fp64": {k: cmp_f[k] for k in ("class_agree", "err_le_1e-6", "attempts_equal")}},
        "mixed_vs_fp64": {k: cmp[k] for k in ("class_agree", "class_agree_outside_band", "err_le_1e-6",
                                             "err_le_1e-6_outside_band", "attempts_equal", "psnr_intensity",
                                             "state_bit_identical")},
        "work": {"attempts": st_m["attempts"], "rhs_evals": st_m["rhs_evals"], "rhs_evals_fp32": st_m["rhs_evals_fp32"],
                 "rhs_evals_fast_fp64": st_m["rhs_evals_fast"],
                 "rhs_evals_exact_fp64": st_m["rhs_evals"] - st_m["rhs_evals_fp32"] - st_m["rhs_evals_fast"],
                 "attempts_per_ray": st_m["attempts"] / n},
        "cpu_baseline": None,
    }
    line["first_frame_value"] = n / (tile_ms["mixed"] * 1e-3)
    line["first_frame_note"] = ("value is a repeat frame of the same scene under the ray order learned from its "
                                "warm-up; first_frame_value is the first frame in 8x8-tile order (kernel time)")
    line["constants_of_motion_drift"] = {p: com_drift(rend, scene, p) for p in ("fp64", "mixed")}
    if not args.no_cpu:
        line["cpu_baseline"], ref = cpu_baseline(scene)
        # every ray of the workload against the FP64 oracle
        shape = (scene.height, scene.width)
        keys = ("class_agree", "class_agree_outside_band", "err_le_1e-6", "err_le_1e-6_outside_band", "err_median",
                "attempts_equal", "psnr_intensity", "state_bit_identical")
        line["fp64_vs_oracle"] = {k: v for k, v in compare(hd, ref, shape=shape).items() if k in keys}
        line["mixed_vs_oracle"] = {k: v for k, v in compare(hm, ref, shape=shape).items() if k in keys}
    if not args.headline_only:
        # the other workloads' frames and the adjoint, each against the oracle
        line["workloads"] = {"kerr4096": frame_kerr4096(torch, rend), "batch64": frame_batch64(torch, rend),
                             "adjoint": block_adjoint(torch, rend)}
        line["adjoint_grads_per_s"] = line["workloads"]["adjoint"]["grads_per_s"]
    rend.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
