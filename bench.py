#!/usr/bin/env python3
"""Benchmark of the GRay2 geodesic hot path on one MI355X.

Headline workload (one "step" = one full render): analytic Kerr a=0.9375,
camera at r=1000 M, inclination 60 deg, 1024x1024 rays over a 64 M x 64 M
image plane, radiative transfer on through the synthetic Keplerian disk with
4 channels nu = {1, 2.67, 4, 8}, dt_dump = -32, n_sub = 1024, 128 dump
intervals, rays stopped at r < 1.01 r_H or r > 1000 M moving outward.

`value` is rays*RK4-steps/s of the FP32 form (persistent ray pool, FP32
bulk, compensated state, FP64 initial conditions and FP64 re-integration of
collapsed-step rays), over the whole render with inputs resident on the GPU.
The plain FP64 form (bitwise equal to the FP64 oracle) is timed on the same
workload and the FP32 image is compared with it (relative L1 per channel).

Further sections (each can be skipped): the numerical spacetime (512^2 rays
through a 128^3 boosted Kerr-Schild grid with two time levels, FP32 form on
an FP32 grid and FP64 form on an FP64 grid, HBM roofline), a sample of the
serving stream (Poisson arrivals of uneven requests through one pool, p50/p99
latency), the CPU baseline, and (with --precision-form) the a=0.998, i=85,
2048^2 precision comparison.
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
PEAK_HBM_GBPS = 8000.0
# HBM bytes per launch of the dominant kernels on these workloads, from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs, scaled by what the same
# counters read for a 256 MiB device copy (read x1.9997, write x1.0000):
# profiles/kerr_disk_1024x1024_kernels.md (k_pool_f32<4>, per pass) and
# profiles/grid_bks128_512x512_kernels.md (k_grid_tiles_f32<0, float> and
# k_grid_tiles_f64<0, double>).
TRAFFIC_POOL_PASS = (64.10 + 80.41) * 2**20
TRAFFIC_GRID_TILES = (525.41 + 21.40) * 2**20
TRAFFIC_GRID_TILES64 = (4012.52 + 21.32) * 2**20
TRAFFIC_SOURCE = "rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE (profiles/)"
# Algorithmic bytes per gridded ray-step: 4 RHS x 40 Christoffels x 8
# corners x 2 time levels x bytes per stored real.
GRID_B = {4: 4 * 40 * 8 * 2 * 4, 8: 4 * 40 * 8 * 2 * 8}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def render(G, precision, w=W, a=A_SPIN, inc=I_OBS):
    r = G.Render(w, w, a, precision=precision, nu=NUS)
    t0 = time.perf_counter()
    r.icond(IMG, IMG, R_OBS, inc, J_OBS)
    ns = r.render(DT, N_SUB, N_DUMP)
    wall = time.perf_counter() - t0
    return r, wall, ns


def l1(img, ref):
    return (np.abs(img - ref).sum(axis=0) / np.abs(ref).sum(axis=0)).tolist()


def headline(G, args, out):
    for _ in range(args.warmup):
        r, _, _ = render(G, G.FP32)
        r.close()
    walls, total_steps, kern_ns, pool_steps = [], 0, 0.0, 0
    img32 = info32 = None
    t_all = time.perf_counter()
    for k in range(args.steps):
        r, wall, ns = render(G, G.FP32)
        walls.append(wall)
        st, nfb = r.stats()
        ps, fbs = r.pass_steps()
        kn, nl = r.kernel_time()
        total_steps += st
        kern_ns += kn
        pool_steps += ps
        if k == args.steps - 1:
            d32, info32 = r.read_f64()
            img32 = d32[:, 8:12].copy()
            out["fallback_rays"] = nfb
            out["fallback_ray_steps"] = fbs
        r.close()
    t_all = time.perf_counter() - t_all
    value = total_steps / t_all
    out.update({
        "value": value,
        "ms_per_step": 1e3 * t_all / args.steps,
        "ray_steps_per_render": total_steps // args.steps,
        "render_latency_ms": {"p50": 1e3 * float(np.percentile(walls, 50)),
                              "p99": 1e3 * float(np.percentile(walls, 99))},
    })
    k_s = kern_ns * 1e-9 / args.steps
    achieved = (pool_steps / args.steps) * FLOP_PER_STEP / k_s / 1e12
    out["roofline"] = {
        "bound": "valu", "kernel": "k_pool_f32<4>",
        "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP32_TFLOPS,
        "frac_of_unpacked_fp32": achieved / PEAK_FP64_TFLOPS,
        "traffic": TRAFFIC_POOL_PASS, "traffic_unit": "B per pass",
        "traffic_source": TRAFFIC_SOURCE,
        "flop_per_ray_step": FLOP_PER_STEP,
        "kernel_ms": 1e3 * k_s,
    }
    return img32, info32, value


def fp64_section(G, img32, info32, value, out):
    r, wall, ns = render(G, G.FP64)
    st64, _ = r.stats()
    d64, info64 = r.read_f64()
    kn64, nl = r.kernel_time()
    r.close()
    img64 = d64[:, 8:12]
    fp64_rate = st64 / wall
    a64 = st64 * FLOP_PER_STEP / (kn64 * 1e-9) / 1e12
    out["fp64"] = {
        "value": fp64_rate, "unit": "ray-steps/s", "render_ms": 1e3 * wall,
        "ray_steps": st64, "evolve_launches": nl,
        "roofline": {"bound": "valu", "kernel": "k_evolve_f64<4>", "achieved": a64,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": a64 / PEAK_FP64_TFLOPS},
    }
    out["speedup_fp32_vs_fp64"] = value / fp64_rate
    out["image_rel_l1_vs_fp64"] = l1(img32, img64)
    out["image_peak_err_vs_fp64"] = float(np.abs(img32 - img64).max() / np.abs(img64).max())
    # null constraint |g(k,k)| at the end, over the rays that escaped (r > 999 M)
    esc = np.sqrt(np.sum(d64[:, 1:4] ** 2, axis=1)) > 999.0
    out["null_drift_max"] = {"fp32": float(np.abs(info32[esc]).max()),
                             "fp64": float(np.abs(info64[esc]).max()),
                             "rays": int(esc.sum())}
    return fp64_rate


def grid_section(G, out):
    from gray_amd import spacetime as S
    n = 128
    xs = S.axis(-40.0, 40.0, n)
    b = S.bks(1.0, 0.7, 0.3, (1, 0, 0))
    lv = [S.bks_sample(b, t, xs, xs, xs) for t in (0.0, 1.0)]
    res = {"workload": "512x512 rays, r_obs=35, i=45, 32Mx32M plane, boosted Kerr-Schild "
                       "a=0.7 v=0.3 along x on 128^3 over [-40,40]^3, levels 0.0 and 1.0, "
                       "t_init=1.0, dt=-32 x 4 dumps, n_sub=1024, stop at box exit"}
    finals = {}
    for dtype, prec in ((4, G.FP32), (8, G.FP64)):
        tp = np.float32 if dtype == 4 else np.float64
        st = G.Spacetime((n,) * 3, (-40.0,) * 3, (40.0,) * 3, dtype=dtype)
        for slot, (g, m) in enumerate(lv):
            st.upload(slot, float(slot), g.astype(tp), m.astype(tp))
        st.set_frozen(False)
        with G.Render(512, 512, precision=prec, spacetime=st, t_init=1.0) as r:
            t0 = time.perf_counter()
            r.icond(32.0, 32.0, 35.0, 45.0, 0.0)
            r.render(-32.0, 1024, 4)
            wall = time.perf_counter() - t0
            kn, nl = r.kernel_time()
            steps = int(r.steps().sum())
            finals[dtype] = (r.read_f64()[0], r.steps())
        st.close()
        gbps = steps * GRID_B[dtype] / (kn * 1e-9) / 1e9
        res[f"{'fp32' if dtype == 4 else 'fp64'}_grid"] = {
            "form": "fp32" if prec == G.FP32 else "fp64", "value": steps / wall,
            "unit": "ray-steps/s", "ray_steps": steps, "render_ms": 1e3 * wall,
            "roofline": {"bound": "hbm", "kernel": "k_grid_tiles_f32<0, float>" if dtype == 4 else "k_grid_tiles_f64<0, double>",
                         "achieved": gbps, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                         "frac": gbps / PEAK_HBM_GBPS,
                         "traffic": TRAFFIC_GRID_TILES if dtype == 4 else TRAFFIC_GRID_TILES64,
                         "traffic_source": TRAFFIC_SOURCE,
                         "bytes_per_ray_step": GRID_B[dtype]},
        }
    (d32, s32), (d64, s64) = finals[4], finals[8]
    same = s32 == s64
    dx = np.abs(d32[same, 1:4] - d64[same, 1:4]).max(axis=1)
    res["fp32_vs_fp64"] = {"same_step_count": float(same.mean()),
                           "median_pos_err": float(np.median(dx)), "max_pos_err": float(dx.max())}
    res["speedup_fp32_vs_fp64"] = res["fp32_grid"]["value"] / res["fp64_grid"]["value"]
    out["grid"] = res


def serving_section(G, args, capacity, cap_name, out):
    from gray_amd import serve as SV
    # Mean work of the sample's requests: analytic Kerr at ~6.4e4 steps/ray,
    # gridded at ~1e4 steps/ray; arrivals come at 1.2x the rate the given
    # capacity (ray-steps/s) sustains.
    o = SV.defaults(n_requests=args.serve_n, max_side=args.serve_max_side)
    items = SV.stream(o)
    est_steps = np.mean([i.side * i.side * (1e4 if i.grid >= 0 else 6.4e4) for i in items])
    o.rate = 1.2 * capacity / est_steps
    rep, lat = SV.run(o)
    res = {
        "sample": f"first {args.serve_n} requests of the seeded stream (seed 0x47524159, "
                  f"image side capped at {args.serve_max_side}), Poisson rate "
                  f"{o.rate:.3f}/s = 1.2x the measured {cap_name} capacity",
        "p50_ms": rep["p50_ms"], "p99_ms": rep["p99_ms"], "max_ms": rep["max_ms"],
        "wall_s": rep["wall_s"], "rays": rep["rays"], "ray_steps": rep["ray_steps"],
        "epochs": rep["epochs"], "value": rep["ray_steps_per_s"], "unit": "ray-steps/s",
    }
    # the same requests arriving together: the pool's saturated throughput
    o.rate = 1e9
    sat, _ = SV.run(o)
    res["saturated"] = {"value": sat["ray_steps_per_s"], "unit": "ray-steps/s",
                        "wall_s": sat["wall_s"], "p50_ms": sat["p50_ms"], "p99_ms": sat["p99_ms"],
                        "epochs": sat["epochs"]}
    # plain FP64 form, run serially on the first requests of the same stream
    t_fp64, st_fp64 = 0.0, 0
    for it in items[: args.serve_fp64_n]:
        if it.grid >= 0:
            continue
        nus = NUS[: it.n_freq]
        r = G.Render(it.side, it.side, it.a_spin, precision=G.FP64, nu=nus)
        t0 = time.perf_counter()
        r.icond(IMG, IMG, R_OBS, it.i_obs, 0.0)
        r.render(DT, N_SUB, N_DUMP)
        t_fp64 += time.perf_counter() - t0
        st_fp64 += r.stats()[0]
        r.close()
    if st_fp64:
        res["fp64_serial"] = {"value": st_fp64 / t_fp64, "unit": "ray-steps/s",
                              "requests": f"analytic-Kerr requests among the first "
                                          f"{args.serve_fp64_n}"}
        res["speedup_saturated_vs_fp64_serial"] = sat["ray_steps_per_s"] / (st_fp64 / t_fp64)
    out["serving"] = res
    out["latency_ms"] = {"p50": rep["p50_ms"], "p99": rep["p99_ms"],
                         "source": "serving sample"}


def precision_section(G, out):
    ra, w32, _ = render(G, G.FP32, w=2048, a=0.998, inc=85.0)
    d32, i32 = ra.read_f64()
    st32, nfb = ra.stats()
    ra.close()
    rb, w64, _ = render(G, G.FP64, w=2048, a=0.998, inc=85.0)
    d64, i64 = rb.read_f64()
    st64, _ = rb.stats()
    rb.close()
    esc = np.sqrt(np.sum(d64[:, 1:4] ** 2, axis=1)) > 999.0
    out["precision_form"] = {
        "workload": "a=0.998, i=85, 2048x2048 rays, disk transfer 4 channels",
        "fp32_value": st32 / w32, "fp64_value": st64 / w64, "speedup": (st32 / w32) / (st64 / w64),
        "image_rel_l1": l1(d32[:, 8:12], d64[:, 8:12]),
        "peak_pixel_err": float(np.abs(d32[:, 8:12] - d64[:, 8:12]).max() / np.abs(d64[:, 8:12]).max()),
        "null_drift_fp32": float(np.abs(i32[esc]).max()), "null_drift_fp64": float(np.abs(i64[esc]).max()),
        "fallback_rays": nfb,
    }


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(stride, n_threads):
    """FP64 oracle (CPU restatement of the reference path) on every
    stride-th row and column of the same image."""
    from oracle import oracle as O
    p = O.make_params(A_SPIN, nu=NUS)
    w = h = W // stride
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
    ap.add_argument("--no-grid", action="store_true", help="skip the numerical-spacetime section")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving sample")
    ap.add_argument("--precision-form", action="store_true",
                    help="also run the a=0.998, i=85, 2048^2 precision comparison (~1 min)")
    ap.add_argument("--serve-n", type=int, default=48)
    ap.add_argument("--serve-max-side", type=int, default=512)
    ap.add_argument("--serve-fp64-n", type=int, default=12)
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
        print("bench.py: runs on one GPU only (rays would shard across GPUs with no "
              "exchange; multi-GPU runs are not implemented)", file=sys.stderr)
        sys.exit(2)

    from gray_amd import grayhip as G
    if G.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    out = {
        "metric": METRIC, "value": None, "unit": "ray-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "kerr-disk-rt: a=0.9375, i=60, r_obs=1000, 1024x1024 rays, "
                        "64Mx64M plane, 4 channels, dt=-32 x 128 dumps, n_sub=1024",
            "rays": W * H, "n_freq": len(NUS), "form": "fp32 pool + fp64 fallback",
        },
    }
    img32, info32, value = headline(G, args, out)
    log(f"headline: {value:.3e} ray-steps/s")
    if not args.no_fp64:
        fp64_section(G, img32, info32, value, out)
        log(f"fp64: {out['fp64']['value']:.3e} ray-steps/s")
    def optional(name, fn):
        # a failing side section is reported in the line instead of losing it
        try:
            fn()
        except Exception as e:  # noqa: BLE001
            out[f"{name}_error"] = repr(e)
            log(f"{name} failed: {e!r}")

    if not args.no_grid:
        def grid():
            grid_section(G, out)
            log(f"grid: fp32 {out['grid']['fp32_grid']['value']:.3e} ray-steps/s")
        optional("grid", grid)
    if not args.no_serve:
        def serve():
            if "fp64" in out:
                serving_section(G, args, out["fp64"]["value"], "plain-FP64", out)
            else:
                serving_section(G, args, value, "FP32-form", out)
            log(f"serving: p50 {out['serving']['p50_ms']:.0f} ms, "
                f"p99 {out['serving']['p99_ms']:.0f} ms")
        optional("serving", serve)
    if args.precision_form:
        def prec():
            precision_section(G, out)
            log("precision form done")
        optional("precision_form", prec)
    if not args.no_cpu:
        def cpu():
            n_thr = min(os.cpu_count() or 1, 16)
            rate, steps, dt, n = cpu_baseline(args.cpu_stride, n_thr)
            out["cpu_baseline"] = {
                "value": rate, "unit": "ray-steps/s", "cores": n_thr, "kind": "port",
                "sample": f"FP64 oracle (CPU restatement of the reference FP64 path, C, "
                          f"-O3 -ffp-contract=off, OpenMP over rays) on every {args.cpu_stride}th "
                          f"row and column ({n} rays, {steps} ray-steps, {dt:.1f} s) on "
                          f"{n_thr} threads of {cpu_model()}",
            }
        optional("cpu_baseline", cpu)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
