#!/usr/bin/env python3
"""Benchmark of the GRay2 geodesic hot path on one MI355X.

Workload (one "step" = one full render): analytic Kerr a=0.9375, camera at
r=1000 M, inclination 60 deg, 1024x1024 rays over a 64 M x 64 M image plane,
radiative transfer on through the synthetic Keplerian disk with 4 channels
nu = {1, 2.67, 4, 8}, dt_dump = -32, n_sub = 1024, 128 dump intervals, rays
stopped at r < 1.01 r_H or r > 1000 M moving outward.

`value` is rays*RK4-steps/s of the FP32 form (persistent ray pool, FP32
bulk, compensated state, FP64 initial conditions and FP64 re-integration of
collapsed-step rays), over the whole render with inputs resident on the GPU.
The plain FP64 form (bitwise equal to the FP64 oracle) is timed on the same
workload and the FP32 image is compared with it (relative L1 per channel).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("rays·RK4 steps/s FP32 vs FP64 at image L1 err≤1e-4; "
          "p50/p99 render latency")

A_SPIN = 0.9375
NUS = (1.0, 2.67, 4.0, 8.0)
W = H = 1024
IMG = 64.0
R_OBS = 1000.0
I_OBS, J_OBS = 60.0, 0.0
DT, N_SUB, N_DUMP = -32.0, 1024, 128

# Algorithmic FLOP per ray-step (operation counts of the reference kernels):
# 4 x 205 (Kerr-Schild gr_rhs) + 96 (RK4 combination) + 16 (step rule), and
# per RHS 100 + 15 n_freq for transfer plus 24 n_freq for the RK4 update.
F_KS = 4 * 205 + 96 + 16


def f_rt(n_freq):
    return 4 * (100 + 15 * n_freq) + 24 * n_freq


FLOP_PER_STEP = F_KS + f_rt(len(NUS))
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6


def render(G, precision):
    r = G.Render(W, H, A_SPIN, precision=precision, nu=NUS)
    t0 = time.perf_counter()
    r.icond(IMG, IMG, R_OBS, I_OBS, J_OBS)
    ns = r.render(DT, N_SUB, N_DUMP)
    wall = time.perf_counter() - t0
    return r, wall, ns


def cpu_baseline(stride, n_threads):
    """FP64 oracle (CPU restatement of the reference path) on every
    stride-th row and column of the same image."""
    from oracle import oracle as O
    p = O.make_params(A_SPIN, nu=NUS)
    w = h = W // stride
    # the strided sub-image: pixel centres of the full image at rows/cols
    # 0, stride, 2*stride, ...
    data, _ = O.icond_all(p, W, H, IMG, IMG, R_OBS, I_OBS, J_OBS)
    sub = np.ascontiguousarray(data.reshape(H, W, -1)[::stride, ::stride].reshape(w * h, -1))
    t0 = time.perf_counter()
    steps = O.evolve_subset(p, sub, 1, DT, N_SUB, N_DUMP, n_threads)
    dt = time.perf_counter() - t0
    return steps / dt, steps, dt, w * h


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 reference run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-stride", type=int, default=16)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit")
    args = ap.parse_args()

    if args.copy_probe:
        import torch
        n = args.copy_probe << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n}), flush=True)
        return

    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (rays shard across GPUs with no "
              "exchange; multi-GPU runs are not part of it)", file=sys.stderr)
        sys.exit(2)

    from gray_amd import grayhip as G
    if G.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    # warmup (untimed)
    for _ in range(args.warmup):
        r, _, _ = render(G, G.FP32)
        r.close()

    walls, total_steps, kern_ns, pool_steps, fb_rays = [], 0, 0.0, 0, 0
    img32 = None
    t_all = time.perf_counter()
    for k in range(args.steps):
        r, wall, ns = render(G, G.FP32)
        walls.append(wall)
        st, nfb = r.stats()
        ps, fs = r.pass_steps()
        kn, _ = r.kernel_time()
        total_steps += st
        kern_ns += kn
        pool_steps += ps
        fb_rays = nfb
        if k == args.steps - 1:
            d32, _ = r.read_f64()
            img32 = d32[:, 8:12].copy()
        r.close()
    t_all = time.perf_counter() - t_all
    value = total_steps / t_all

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "ray-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_all / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "kerr-disk-rt: a=0.9375, i=60, r_obs=1000, 1024x1024 rays, "
                        "64Mx64M plane, 4 channels, dt=-32 x 128 dumps, n_sub=1024",
            "rays": W * H, "n_freq": len(NUS), "form": "fp32 pool + fp64 fallback",
        },
        "ray_steps_per_render": total_steps // args.steps,
        "fallback_rays": fb_rays,
        "latency_ms": {"p50": 1e3 * float(np.percentile(walls, 50)),
                       "p99": 1e3 * float(np.percentile(walls, 99))},
    }

    # dominant kernel: the FP32 ray pool
    k_s = kern_ns * 1e-9 / args.steps
    achieved = (pool_steps / args.steps) * FLOP_PER_STEP / k_s / 1e12
    out["roofline"] = {
        "bound": "valu", "kernel": "k_pool_f32<4>",
        "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP32_TFLOPS,
        "frac_of_unpacked_fp32": achieved / PEAK_FP64_TFLOPS,
        "traffic": None,
        "flop_per_ray_step": FLOP_PER_STEP,
        "kernel_ms": 1e3 * k_s,
    }

    if not args.no_fp64:
        r, wall, ns = render(G, G.FP64)
        st64, _ = r.stats()
        d64, _ = r.read_f64()
        kn64, nl = r.kernel_time()
        r.close()
        img64 = d64[:, 8:12]
        l1 = (np.abs(img32 - img64).sum(axis=0) / np.abs(img64).sum(axis=0)).tolist()
        peak_err = float(np.abs(img32 - img64).max() / np.abs(img64).max())
        fp64_rate = st64 / wall
        a64 = st64 * FLOP_PER_STEP / (kn64 * 1e-9) / 1e12
        out["fp64"] = {
            "value": fp64_rate, "unit": "ray-steps/s", "render_ms": 1e3 * wall,
            "ray_steps": st64, "evolve_launches": nl,
            "roofline": {"bound": "valu", "achieved": a64, "peak": PEAK_FP64_TFLOPS,
                         "unit": "TFLOP/s", "frac": a64 / PEAK_FP64_TFLOPS},
        }
        out["speedup_fp32_vs_fp64"] = value / fp64_rate
        out["image_rel_l1_vs_fp64"] = l1
        out["image_peak_err_vs_fp64"] = peak_err

    if not args.no_cpu:
        n_thr = min(os.cpu_count() or 1, 16)
        rate, steps, dt, n = cpu_baseline(args.cpu_stride, n_thr)
        out["cpu_baseline"] = {
            "value": rate, "unit": "ray-steps/s", "cores": n_thr, "kind": "port",
            "sample": f"FP64 oracle (CPU restatement of the reference path) on every "
                      f"{args.cpu_stride}th row and column ({n} rays, {steps} ray-steps, "
                      f"{dt:.1f} s)",
        }

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
