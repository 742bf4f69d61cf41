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
serving stream (a stratified sample of uneven requests through one pool:
p50/p99 latency, saturated throughput, and the same requests rendered alone
one after another in the plain FP64 form and in the FP32 form), the a=0.998,
i=85, 2048^2 precision comparison, and the CPU baseline.
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
# Hardware FLOP per ray-step of the kernels themselves on the headline
# workload: rocprofv3 SQ_INSTS_VALU_FLOPS_FP32 (FP64) x 64 lanes / ray-steps
# (FMA = 2; profiles/kerr_disk_1024x1024_kernels.md).  The FP32 form
# integrates the Hamiltonian form of the geodesic, which needs fewer
# operations per ray-step than the reference's u^mu form that FLOP_PER_STEP
# counts; the plain FP64 form evaluates the reference's form, plus the
# Newton steps of its correctly rounded divisions and square roots.
HW_FLOP_POOL = 799
HW_FLOP_FP64 = 2176
HW_SOURCE = "rocprofv3 --pmc SQ_INSTS_VALU_FLOPS_FP32/FP64 x 64 / ray-steps (profiles/)"
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBPS = 8000.0
# HBM bytes per launch of the dominant kernels on these workloads, from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs, scaled by what the same
# counters read for a 256 MiB device copy (read x1.9997, write x1.0000):
# profiles/kerr_disk_1024x1024_kernels.md (k_pool_f32<4>, per pass) and
# profiles/grid_bks128_512x512_kernels.md (k_grid_tiles_f32<0, float> and
# k_grid_tiles_f64<0, double>).
TRAFFIC_POOL_PASS = (57.80 + 82.81) * 2**20
TRAFFIC_GRID_TILES = (526.93 + 21.78) * 2**20
TRAFFIC_GRID_TILES64 = (4695.65 + 22.24) * 2**20
TRAFFIC_SOURCE = "rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE (profiles/)"
# Gridded RHS: 8 stencil corners x 40 Christoffels of one time level, read
# from the wave's LDS brick (or L1/L2 for a lane outside it).  Both levels
# are read only while the ray's grid time lies between them: on the bench
# workload (t_init = 1, levels 0 and 1, u^t = 1 in the coordinate-time
# form, step 0.01) that is a ray's first 100 RK4 steps.
GRID_CORNER_B = {4: 8 * 40 * 4, 8: 8 * 40 * 8}
GRID_BLEND_STEPS = 100
# LDS read bandwidth: 256 CUs x 128 B/clk x 2.4 GHz.
PEAK_LDS_GBPS = 256 * 128 * 2.4


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


_FMA_PROBE = None


def fma_probe():
    """Attainable non-packed FMA rates of this GPU (gray-fma-probe: 8
    independent v_fma chains per lane, 16 waves per SIMD), the ceiling of a
    VALU kernel built without packed FP32; measured once per run."""
    global _FMA_PROBE
    if _FMA_PROBE is None:
        import subprocess
        exe = os.path.join(ROOT, "gray_amd", "gray-fma-probe")
        try:
            r = subprocess.run([exe], capture_output=True, text=True, timeout=120, check=True)
            _FMA_PROBE = json.loads(r.stdout.strip().splitlines()[-1])
        except (OSError, subprocess.SubprocessError, ValueError, IndexError) as e:
            _FMA_PROBE = {"error": repr(e)}
    return _FMA_PROBE


def hw_roofline(achieved, flop, peak, key, key_occ):
    """The kernel's own arithmetic rate (hardware FLOP counted by rocprofv3)
    against the peak and against the FMA probe."""
    return {"flop_per_ray_step": flop, "source": HW_SOURCE, "achieved": achieved,
            "unit": "TFLOP/s", "frac": achieved / peak, **probe_fracs(achieved, key, key_occ)}


def probe_fracs(achieved, key, key_occ):
    """Fractions of the probe's FMA rate at 16 waves per SIMD and at the
    kernel's own occupancy."""
    p = fma_probe()
    if key not in p or key_occ not in p:
        return {"fma_probe": p}
    return {"fma_probe_tflops": p[key], "frac_of_fma_probe": achieved / p[key],
            "fma_probe_tflops_at_occupancy": p[key_occ],
            "frac_of_fma_probe_at_occupancy": achieved / p[key_occ]}


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
    """Renders of the 1024^2 workload on one context (its device buffers
    allocated once, as a server keeps them): each step sets up the camera
    rays on the device and renders them."""
    r = G.Render(W, W, A_SPIN, precision=G.FP32, nu=NUS)
    for _ in range(args.warmup):
        r.icond(IMG, IMG, R_OBS, I_OBS, J_OBS)
        r.render(DT, N_SUB, N_DUMP)
    walls, total_steps, kern_ns, pool_steps = [], 0, 0.0, 0
    img32 = info32 = None
    t_all = time.perf_counter()
    for k in range(args.steps):
        t0 = time.perf_counter()
        r.icond(IMG, IMG, R_OBS, I_OBS, J_OBS)
        r.render(DT, N_SUB, N_DUMP)
        walls.append(time.perf_counter() - t0)
        st, nfb = r.stats()
        ps, fbs = r.pass_steps()
        kn, nl = r.kernel_time()
        total_steps += st
        kern_ns += kn
        pool_steps += ps
        if k == args.steps - 1:
            out["fallback_rays"] = nfb
            out["fallback_ray_steps"] = fbs
    t_all = time.perf_counter() - t_all
    d32, info32 = r.read_f64()
    img32 = d32[:, 8:12].copy()
    r.close()
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
        "traffic": TRAFFIC_POOL_PASS, "traffic_unit": "B per pass",
        "traffic_source": TRAFFIC_SOURCE,
        "flop_per_ray_step": FLOP_PER_STEP,
        "flop_per_ray_step_is": "the reference's u^mu-form operation count per ray-step",
        "kernel_ms": 1e3 * k_s,
        "kernel_ms_is": "span from the first pool pass's start to the last pass's end "
                        "(the passes overlap on two streams); per-dispatch durations are in "
                        "the rocprofv3 kernel trace under profiles/",
        "hw": hw_roofline((pool_steps / args.steps) * HW_FLOP_POOL / k_s / 1e12, HW_FLOP_POOL,
                          PEAK_FP32_TFLOPS, "fp32_fma_tflops", "fp32_fma_tflops_4_waves"),
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
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": a64 / PEAK_FP64_TFLOPS,
                     "kernel_ms": kn64 * 1e-6, "kernel_ms_is": "sum of the evolve launches",
                     "hw": hw_roofline(st64 * HW_FLOP_FP64 / (kn64 * 1e-9) / 1e12, HW_FLOP_FP64,
                                       PEAK_FP64_TFLOPS, "fp64_fma_tflops",
                                       "fp64_fma_tflops_2_waves")},
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
    lv = [S.bks_sample(b, t, xs, xs, xs, rho=True) for t in (0.0, 1.0)]
    res = {"workload": "512x512 rays, r_obs=35, i=45, 32Mx32M plane, boosted Kerr-Schild "
                       "a=0.7 v=0.3 along x on 128^3 over [-40,40]^3, levels 0.0 and 1.0, "
                       "t_init=1.0, dt=-32 x 4 dumps, n_sub=1024, stop at box exit"}
    finals = {}
    for dtype, prec in ((4, G.FP32), (8, G.FP64)):
        tp = np.float32 if dtype == 4 else np.float64
        st = G.Spacetime((n,) * 3, (-40.0,) * 3, (40.0,) * 3, dtype=dtype)
        for slot, (g, m, _) in enumerate(lv):
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
        s_ray = finals[dtype][1]
        levels = 1.0 + float(np.minimum(s_ray, GRID_BLEND_STEPS).sum()) / steps
        b_step = 4 * GRID_CORNER_B[dtype] * levels
        k_s = kn * 1e-9
        lds = steps * b_step / k_s / 1e9
        traffic = TRAFFIC_GRID_TILES if dtype == 4 else TRAFFIC_GRID_TILES64
        unique = 2 * n ** 3 * 40 * dtype + 512 * 512 * 8 * 8 * 2
        hbm = traffic / k_s / 1e9
        roof = {
            "kernel": "k_grid_tiles_f32<0, float>" if dtype == 4 else "k_grid_tiles_f64<0, double>",
            "kernel_ms": 1e3 * k_s,
            "lds": {"achieved": lds, "peak": PEAK_LDS_GBPS, "unit": "GB/s",
                    "frac": lds / PEAK_LDS_GBPS, "bytes_per_ray_step": b_step,
                    "levels_read_per_rhs": levels,
                    "model": "4 RHS x 8 corners x 40 Christoffels x sizeof(T) x levels read"},
            "hbm": {"achieved": hbm, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                    "frac": hbm / PEAK_HBM_GBPS, "traffic": traffic, "unique_bytes": unique,
                    "traffic_over_unique": traffic / unique, "traffic_source": TRAFFIC_SOURCE,
                    "unique_model": "two time levels x n^3 x 40 Christoffels x sizeof(T) + "
                                    "ray state read and written"},
        }
        # Both kernels read their corners from LDS bricks and LDS is their
        # busiest unit (FP32 0.59, FP64 0.46 plus bank conflicts on 23% of
        # CU cycles; VALU issue 0.56 and 0.50).  HBM is idle: the L2 misses
        # come to a few GB/s, well under 1% of 8 TB/s.
        bound = "lds"
        roof.update({"bound": bound, **{k: roof[bound][k] for k in ("achieved", "peak", "unit", "frac")}})
        res[f"{'fp32' if dtype == 4 else 'fp64'}_grid"] = {
            "form": "fp32" if prec == G.FP32 else "fp64", "value": steps / wall,
            "unit": "ray-steps/s", "ray_steps": steps, "render_ms": 1e3 * wall,
            "roofline": roof,
        }
    (d32, s32), (d64, s64) = finals[4], finals[8]
    same = s32 == s64
    dx = np.abs(d32[same, 1:4] - d64[same, 1:4]).max(axis=1)
    res["fp32_vs_fp64"] = {"same_step_count": float(same.mean()),
                           "median_pos_err": float(np.median(dx)),
                           "p99_pos_err": float(np.percentile(dx, 99)),
                           "max_pos_err": float(dx.max())}
    res["speedup_fp32_vs_fp64"] = res["fp32_grid"]["value"] / res["fp64_grid"]["value"]
    res["transfer"] = grid_transfer(G, lv)
    out["grid"] = res


def grid_transfer(G, lv, nus=(2.3e11, 3.45e11, 5.0e11, 8.6e11)):
    """The same rays and grid with radiative transfer: 4 channels of the
    reference flow (thermal bremsstrahlung and Kirchhoff absorption on the
    generator's density 1/|x|, capped at 100 as the tests do), FP32 form on
    the FP32 grid against the plain FP64 form on the FP64 grid."""
    n = 128
    res = {"channels_hz": list(nus), "fluid": "thermal (flow.cl, b = 0)"}
    imgs = {}
    for dtype, prec in ((4, G.FP32), (8, G.FP64)):
        tp = np.float32 if dtype == 4 else np.float64
        st = G.Spacetime((n,) * 3, (-40.0,) * 3, (40.0,) * 3, dtype=dtype)
        for slot, (g, m, rho) in enumerate(lv):
            st.upload(slot, float(slot), g.astype(tp), m.astype(tp))
            st.upload_rho(slot, np.minimum(rho, 100.0).astype(tp))
        st.set_frozen(False)
        with G.Render(512, 512, precision=prec, spacetime=st, t_init=1.0, nu=nus) as r:
            t0 = time.perf_counter()
            r.icond(32.0, 32.0, 35.0, 45.0, 0.0)
            r.render(-32.0, 1024, 4)
            wall = time.perf_counter() - t0
            kn, _ = r.kernel_time()
            steps = int(r.steps().sum())
            dfin = r.read_f64()[0]
            imgs[dtype] = dfin[:, 8:8 + len(nus)]
            if dtype == 8:
                # rays that end at the box wall (the others stop at the
                # horizon proxy, where the static flow's shift -k_t can pass
                # through zero and dI ~ 1/shift^2 has no scale)
                leave = np.abs(dfin[:, 1:4]).max(axis=1) > 39.0
        st.close()
        res["fp32" if dtype == 4 else "fp64"] = {
            "kernel": f"k_grid_tiles_{'f32' if dtype == 4 else 'f64'}<4, "
                      f"{'float' if dtype == 4 else 'double'}>",
            "value": steps / wall, "unit": "ray-steps/s", "ray_steps": steps,
            "render_ms": 1e3 * wall, "kernel_ms": kn * 1e-6}
    a, b = imgs[4], imgs[8]
    fin = np.isfinite(a).all(1) & np.isfinite(b).all(1)
    res["image_rel_l1_fp32_vs_fp64"] = (np.abs(a[fin] - b[fin]).sum(0) /
                                        np.abs(b[fin]).sum(0)).tolist()
    res["finite_share"] = float(fin.mean())
    esc = fin & leave
    res["image_rel_l1_leaving_rays"] = (np.abs(a[esc] - b[esc]).sum(0) /
                                        np.abs(b[esc]).sum(0)).tolist()
    res["leaving_share"] = float(leave.mean())
    err = np.abs(a[fin, 0] - b[fin, 0])
    top = np.sort(err)[::-1][:max(1, err.size // 1000)]
    res["error_share_of_worst_0.1pct_rays"] = float(top.sum() / err.sum()) if err.sum() > 0 else 0.0
    res["speedup_fp32_vs_fp64"] = res["fp32"]["value"] / res["fp64"]["value"]
    return res


# Stratified sample of the seeded 2000-request stream (seed 0x47524159):
# the first requests of each resolution in stream order, in proportion to
# the stream's weights, so every size (1024^2 included) and both resident
# grids are served.
SERVE_STRATA = ((64, 10), (128, 10), (256, 6), (512, 4), (1024, 2))


def serve_sample(SV):
    o = SV.defaults()
    want = dict(SERVE_STRATA)
    got = {s: 0 for s in want}
    picked = []
    for it in SV.stream(o):
        if got[it.side] < want[it.side]:
            picked.append(it)
            got[it.side] += 1
    return o, picked


def serve_grids(G, o):
    """The two resident spacetimes of the stream (boosted Kerr-Schild a=0.7,
    v=0.3 along x on uniform [-40,40]^3, levels 0 and 1), as FP32 and FP64
    copies for the single-render forms."""
    from gray_amd import spacetime as S
    b = S.bks(1.0, 0.7, 0.3, (1, 0, 0))
    out = {}
    for k in range(2):
        n = o.grid_n[k]
        xs = S.axis(-40.0, 40.0, n)
        lv = [S.bks_sample(b, float(l), xs, xs, xs, rho=True) for l in range(2)]
        for dtype in (4, 8):
            tp = np.float32 if dtype == 4 else np.float64
            st = G.Spacetime((n,) * 3, (-40.0,) * 3, (40.0,) * 3, dtype=dtype)
            for l, (g, m, rho) in enumerate(lv):
                st.upload(l, float(l), g.astype(tp), m.astype(tp))
                st.upload_rho(l, rho.astype(tp))
            st.set_frozen(False)
            out[(k, dtype)] = st
        del lv
    return out


def render_item(G, it, nf, precision, grids):
    """One request of the stream rendered alone (the request loop's
    parameters and channel frequencies, gray_serve.c fill_request); returns
    (data, ray-steps, wall s)."""
    from gray_amd import serve as SV
    if it.grid >= 0:
        nus = SV.GRID_NUS[:nf]
        st = grids[(it.grid, 4 if precision == G.FP32 else 8)]
        r = G.Render(it.side, it.side, precision=precision, spacetime=st, t_init=1.0, nu=nus)
        cam, sched = (32.0, 32.0, 35.0), (DT, N_SUB, 4)
    else:
        nus = SV.KERR_NUS[:nf]
        r = G.Render(it.side, it.side, it.a_spin, precision=precision, nu=nus)
        cam, sched = (IMG, IMG, R_OBS), (DT, N_SUB, N_DUMP)
    t0 = time.perf_counter()
    r.icond(*cam, it.i_obs, 0.0)
    r.render(*sched)
    wall = time.perf_counter() - t0
    d, _ = r.read_f64()
    steps = int(r.steps().sum())
    r.close()
    return d, steps, wall


def serving_section(G, args, capacity, cap_name, out):
    from gray_amd import serve as SV
    o, items = serve_sample(SV)
    nfs = [SV.served_channels(it) for it in items]
    sizes = {s: sum(1 for it in items if it.side == s) for s, _ in SERVE_STRATA}
    n_grid = sum(1 for it in items if it.grid >= 0)
    # mean work per request (analytic Kerr ~6.4e4 steps/ray, gridded ~1e4):
    # Poisson arrivals at 1.2x the rate the given capacity sustains
    est_steps = np.mean([i.side * i.side * (1e4 if i.grid >= 0 else 6.4e4) for i in items])
    rate = 1.2 * capacity / est_steps
    rng = np.random.default_rng(0x47524159)
    t = np.cumsum(rng.exponential(1.0 / rate, len(items)))
    for it, ta in zip(items, t):
        it.arrival_s = float(ta - t[0])
    rep, lat, res = SV.run_items(o, items, keep_results=True)
    res_ = {
        "sample": f"{len(items)} requests of the seeded stream (seed 0x47524159): the first of "
                  f"each side in stream order, {sizes} ({n_grid} gridded on the 96^3 and 128^3 "
                  f"grids); Poisson arrivals at {rate:.3f}/s = 1.2x the measured {cap_name} "
                  f"capacity",
        "grid_channels": "carried" if SV.GRID_CHANNELS else
                         "none (gridded requests are served without their drawn channels)",
        "p50_ms": rep["p50_ms"], "p99_ms": rep["p99_ms"], "max_ms": rep["max_ms"],
        "wall_s": rep["wall_s"], "rays": rep["rays"], "ray_steps": rep["ray_steps"],
        "epochs": rep["epochs"], "value": rep["ray_steps_per_s"], "unit": "ray-steps/s",
    }
    # the same requests arriving together: the pool's saturated throughput
    for it in items:
        it.arrival_s = 0.0
    sat, _, _ = SV.run_items(o, items)
    served_steps = sum(sat["ray_steps"])
    res_["saturated"] = {"value": sat["ray_steps_per_s"], "unit": "ray-steps/s",
                         "wall_s": sat["wall_s"], "p50_ms": sat["p50_ms"],
                         "p99_ms": sat["p99_ms"], "epochs": sat["epochs"],
                         "ray_steps": served_steps}
    # the same requests, one after another, rendered alone: the plain FP64
    # form (the denominator of the serving target) and the single-render
    # FP32 form (render pool / grid tile kernel)
    grids = serve_grids(G, o)
    try:
        serial = {}
        l1_max, l1_items, fin_min, l1_grid_all = 0.0, 0, 1.0, 0.0
        for prec, name in ((G.FP64, "fp64_serial"), (G.FP32, "fp32_serial")):
            t_all, st_all = 0.0, 0
            for k, it in enumerate(items):
                d, steps, wall = render_item(G, it, nfs[k], prec, grids)
                t_all += wall
                st_all += steps
                if prec == G.FP64 and nfs[k] > 0:
                    # over the rays finite in both images: the reference's
                    # thermal emission makes rays that reach the grid's
                    # central density spike non-finite, in both forms
                    I0, I1 = d[:, 8:8 + nfs[k]], res[k][:, 8:8 + nfs[k]]
                    fin = np.isfinite(I0).all(1) & np.isfinite(I1).all(1)
                    use = fin
                    if it.grid >= 0:
                        # gridded: the rays that leave the box (those stopped
                        # at the horizon proxy carry dI ~ 1/shift^2 with a
                        # shift that can pass through zero; grid.transfer)
                        a0 = np.abs(I1[fin] - I0[fin]).sum(0) / np.abs(I0[fin]).sum(0)
                        l1_grid_all = max(l1_grid_all, float(np.nanmax(a0)))
                        use = fin & (np.abs(d[:, 1:4]).max(axis=1) > 39.0)
                    den = np.abs(I0[use]).sum(0)
                    ok = den > 0
                    if ok.any():
                        err = np.abs(I1[use] - I0[use]).sum(0)[ok] / den[ok]
                        l1_max = max(l1_max, float(err.max()))
                        l1_items += 1
                        fin_min = min(fin_min, float(fin.mean()))
            serial[name] = {"value": st_all / t_all, "unit": "ray-steps/s", "wall_s": t_all,
                            "ray_steps": st_all}
    finally:
        for st in grids.values():
            st.close()
    res_.update(serial)
    res_["fp64_serial"]["requests"] = "the same sample, each rendered alone in the plain FP64 form"
    res_["fp32_serial"]["requests"] = "the same sample, each rendered alone in the FP32 form"
    res_["speedup_saturated_vs_fp64_serial"] = sat["ray_steps_per_s"] / serial["fp64_serial"]["value"]
    res_["saturated_vs_fp32_serial"] = sat["ray_steps_per_s"] / serial["fp32_serial"]["value"]
    res_["image_rel_l1_vs_fp64_max"] = l1_max
    res_["image_rel_l1_requests"] = l1_items
    res_["image_rel_l1_finite_share_min"] = fin_min
    res_["image_rel_l1_rays"] = ("analytic Kerr: every finite ray; gridded: the finite rays "
                                 "that leave the box")
    res_["image_rel_l1_grid_all_finite_rays_max"] = l1_grid_all
    out["serving"] = res_
    out["latency_ms"] = {"p50": rep["p50_ms"], "p99": rep["p99_ms"],
                         "source": "serving sample"}


def precision_section(G, out):
    """The precision workload (a=0.998, i=85, 2048^2 rays, 4 channels): FP32
    form against the plain FP64 form on the whole image."""
    ra, w32, _ = render(G, G.FP32, w=2048, a=0.998, inc=85.0)
    d32, i32 = ra.read_f64()
    st32, nfb = ra.stats()
    ps, fbs = ra.pass_steps()
    kn, _ = ra.kernel_time()
    ra.close()
    rb, w64, _ = render(G, G.FP64, w=2048, a=0.998, inc=85.0)
    d64, i64 = rb.read_f64()
    st64, _ = rb.stats()
    kn64, _ = rb.kernel_time()
    rb.close()
    esc = np.sqrt(np.sum(d64[:, 1:4] ** 2, axis=1)) > 999.0
    a32 = ps * FLOP_PER_STEP / (kn * 1e-9) / 1e12
    a64 = st64 * FLOP_PER_STEP / (kn64 * 1e-9) / 1e12
    out["precision_form"] = {
        "workload": "a=0.998, i=85, r_obs=1000, 2048x2048 rays, 64Mx64M plane, disk transfer "
                    "4 channels, dt=-32 x 128 dumps, n_sub=1024",
        "fp32_value": st32 / w32, "fp64_value": st64 / w64, "speedup": (st32 / w32) / (st64 / w64),
        "fp32_render_ms": 1e3 * w32, "fp64_render_ms": 1e3 * w64,
        "ray_steps": {"fp32_form": st32, "fp64_form": st64, "fp32_pool": ps, "fp64_fallback": fbs},
        "image_rel_l1": l1(d32[:, 8:12], d64[:, 8:12]),
        "peak_pixel_err": float(np.abs(d32[:, 8:12] - d64[:, 8:12]).max() / np.abs(d64[:, 8:12]).max()),
        "null_drift_fp32": float(np.abs(i32[esc]).max()), "null_drift_fp64": float(np.abs(i64[esc]).max()),
        "escaping_rays": int(esc.sum()),
        "fallback_rays": nfb,
        "roofline_fp32_pool": {"kernel": "k_pool_f32<4>", "achieved": a32, "peak": PEAK_FP32_TFLOPS,
                               "unit": "TFLOP/s", "frac": a32 / PEAK_FP32_TFLOPS,
                               "kernel_ms": 1e-6 * kn, "kernel_ms_is": "span of the pool passes",
                               "hw_frac": ps * HW_FLOP_POOL / (kn * 1e-9) / 1e12 / PEAK_FP32_TFLOPS,
                               "hw_frac_is": "hardware FLOP per ray-step of the headline "
                                             "workload"},
        "roofline_fp64": {"kernel": "k_evolve_f64<4>", "achieved": a64, "peak": PEAK_FP64_TFLOPS,
                          "unit": "TFLOP/s", "frac": a64 / PEAK_FP64_TFLOPS},
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


def cpu_threads():
    """Threads for the CPU baseline: the host CPUs this process may use,
    capped by OMP_NUM_THREADS when set (the GPU host's share of CPUs per
    GPU); also returns the host's logical CPU count and the allowed count."""
    nproc = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = nproc
    n = allowed
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n, nproc, allowed


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
    ap.add_argument("--no-headline", action="store_true",
                    help="skip the headline renders (to profile one side section)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 reference run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-grid", action="store_true", help="skip the numerical-spacetime section")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving sample")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the a=0.998, i=85, 2048^2 precision comparison (~40 s)")
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
    value = None
    if not args.no_headline:
        img32, info32, value = headline(G, args, out)
        log(f"headline: {value:.3e} ray-steps/s")
    if not args.no_fp64 and value is not None:
        fp64_section(G, img32, info32, value, out)
        log(f"fp64: {out['fp64']['value']:.3e} ray-steps/s")
    if not args.no_headline:
        out["fma_probe"] = fma_probe()
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
                # without the FP64 run: the plain FP64 form's capacity on the
                # headline workload as measured in full runs (1.58e10 ray-steps/s)
                serving_section(G, args, 1.58e10, "plain-FP64 (earlier full run)", out)
            log(f"serving: p50 {out['serving']['p50_ms']:.0f} ms, "
                f"p99 {out['serving']['p99_ms']:.0f} ms")
        optional("serving", serve)
    if not args.no_precision:
        def prec():
            precision_section(G, out)
            log("precision form done")
        optional("precision_form", prec)
    if not args.no_cpu:
        def cpu():
            n_thr, nproc, allowed = cpu_threads()
            rate, steps, dt, n = cpu_baseline(args.cpu_stride, n_thr)
            out["cpu_baseline"] = {
                "value": rate, "unit": "ray-steps/s", "cores": n_thr, "kind": "port",
                "host_nproc": nproc, "host_cpus_allowed": allowed,
                "per_thread": rate / n_thr,
                "sample": f"FP64 oracle (CPU restatement of the reference FP64 path, C, "
                          f"-O3 -ffp-contract=off, OpenMP over rays) on every {args.cpu_stride}th "
                          f"row and column ({n} rays, {steps} ray-steps, {dt:.1f} s) on "
                          f"{n_thr} threads of {cpu_model()} (host: {nproc} logical CPUs, "
                          f"{allowed} allowed to this process; OMP_NUM_THREADS="
                          f"{os.environ.get('OMP_NUM_THREADS', 'unset')})",
            }
        optional("cpu_baseline", cpu)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
