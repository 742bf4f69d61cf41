"""GPU parity of the numerical-spacetime path (boosted Kerr-Schild
grids at small sizes) against the FP64 grid oracle, through the C ABI.

Bar: the plain FP64 form on uniform grids evaluates the oracle's sequence
and must match it bit for bit (state and step counts), for FP64 and FP32
grid storage (FP32 values are converted exactly).  Fisheye axes go through
cbrt/asinh/sinh, which are not correctly rounded: 1e-10 relative there.
The FP32 form is compared with the FP64 form.
"""
import os
import subprocess

import numpy as np
import pytest

from gray_amd import grayhip as G
from gray_amd import job as J
from gray_amd import spacetime as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BOX = 40.0


def make_grid(n, dtype, times=(0.0,), fisheye=False, a=0.7, v=0.3):
    xs = S.axis(-BOX, BOX, n, fisheye)
    b = S.bks(1.0, a, v, (1, 0, 0))
    lv = [S.bks_sample(b, t, xs, xs, xs) for t in times]
    t = np.float32 if dtype == 4 else np.float64
    lv = [(g.astype(t), m.astype(t)) for g, m in lv]
    return xs, lv


def device_grid(n, dtype, lv, times, fisheye=False):
    st = G.Spacetime((n,) * 3, (-BOX,) * 3, (BOX,) * 3,
                     map=G.MAP_FISHEYE if fisheye else G.MAP_UNIFORM, dtype=dtype)
    for slot, (g, m) in enumerate(lv):
        st.upload(slot, times[slot], g, m)
    st.set_frozen(len(lv) == 1)
    return st


def oracle_grid(n, lv, times, t_off=0.0, fisheye=False):
    return O.GridSpacetime((n,) * 3, (-BOX,) * 3, (BOX,) * 3,
                           [g.astype(np.float64) for g, _ in lv],
                           [m.astype(np.float64) for _, m in lv],
                           times=(times[0], times[-1]), t_off=t_off,
                           map=O.MAP_FISHEYE if fisheye else O.MAP_UNIFORM)


@pytest.mark.parametrize("dtype", [4, 8])
def test_sample_bitwise(dtype):
    n = 17
    xs, lv = make_grid(n, dtype, times=(0.0, 1.0))
    st = device_grid(n, dtype, lv, (0.0, 1.0))
    og = oracle_grid(n, lv, (0.0, 1.0), t_off=0.25)
    rng = np.random.default_rng(2)
    ev = np.column_stack([rng.uniform(-1, 2, 200), rng.uniform(-45, 45, (200, 3))])
    for which, comp in [(0, 0), (0, 13), (0, 39), (1, 0), (1, 9)]:
        got = st.sample(ev, which, comp, t_init=0.25)
        want = np.array([og.sample(which, comp, e) for e in ev])
        assert np.array_equal(got, want), (which, comp, np.abs(got - want).max())


def render_pair(n, dtype, times, w=12, h=12, n_dump=2, precision=G.FP64, fisheye=False,
                t_init=0.0):
    xs, lv = make_grid(n, dtype, times=times, fisheye=fisheye)
    st = device_grid(n, dtype, lv, times, fisheye)
    with G.Render(w, h, precision=precision, spacetime=st, t_init=t_init) as r:
        r.icond(32.0, 32.0, 35.0, 45.0, 0.0)
        for _ in range(n_dump):
            r.evolve(-32.0, 1024)
        d1, i1 = r.read_f64()
        s1 = r.steps()
    og = oracle_grid(n, lv, times, t_off=t_init, fisheye=fisheye)
    d0, i0, s0 = og.render(w, h, 32.0, 32.0, 35.0, 45.0, 0.0, n_dump=n_dump)
    st.close()
    return (d0, i0, s0), (d1, i1, s1)


@pytest.mark.parametrize("dtype", [4, 8])
def test_render_fp64_frozen_bitwise(dtype):
    (d0, i0, s0), (d1, i1, s1) = render_pair(33, dtype, (0.0,))
    assert np.array_equal(s0, s1)
    assert np.array_equal(d0, d1), np.abs(d0 - d1).max()
    assert np.array_equal(i0, i1)


@pytest.mark.parametrize("w,h", [(1, 1), (1, 17), (9, 1), (13, 7)])
def test_render_fp64_ragged_images_bitwise(w, h):
    """Images that are not whole 8x8 tiles (a lone ray, a single column or
    row, tiles cut on both sides) through the tile kernel: bit for bit the
    oracle's."""
    (d0, i0, s0), (d1, i1, s1) = render_pair(33, 8, (0.0, 1.0), w=w, h=h, t_init=1.0)
    assert np.array_equal(s0, s1)
    assert np.array_equal(d0, d1), np.abs(d0 - d1).max()
    assert np.array_equal(i0, i1)


@pytest.mark.parametrize("w,h", [(1, 1), (1, 17), (9, 1), (13, 7)])
def test_render_fp32_ragged_images(w, h):
    """The same shapes in the FP32 form against the plain FP64 form."""
    n, times = 33, (0.0, 1.0)
    xs, lv = make_grid(n, 4, times=times)
    st = device_grid(n, 4, lv, times)
    out = []
    for prec in (G.FP64, G.FP32):
        with G.Render(w, h, precision=prec, spacetime=st, t_init=1.0) as r:
            r.icond(32.0, 32.0, 35.0, 45.0, 0.0)
            r.render(-32.0, 1024, 2)
            out.append((r.read_f64()[0], r.steps()))
    st.close()
    (d64, s64), (d32, s32) = out
    assert d32.shape == d64.shape == (w * h, 8)
    assert np.all(s32 > 0)
    same = s64 == s32
    if w * h >= 8:
        assert same.mean() >= 0.5
    if same.any():
        assert np.abs(d32[same, 1:4] - d64[same, 1:4]).max() < 1e-2


def test_render_fp64_two_levels_bitwise():
    (d0, i0, s0), (d1, i1, s1) = render_pair(33, 4, (0.0, 1.0), t_init=1.0)
    assert np.array_equal(s0, s1)
    assert np.array_equal(d0, d1), np.abs(d0 - d1).max()


def test_render_fp64_fisheye_close():
    (d0, i0, s0), (d1, i1, s1) = render_pair(33, 8, (0.0,), fisheye=True, n_dump=1)
    assert (s0 == s1).mean() > 0.95
    same = s0 == s1
    rel = np.abs(d1[same] - d0[same]) / np.maximum(np.abs(d0[same]), 1e-300)
    assert rel.max() < 1e-9, rel.max()


def test_render_fp32_form_vs_fp64():
    """FP32 form on an FP32 grid against the FP64 form on the FP64 grid."""
    n, times = 33, (0.0,)
    xs, lv64 = make_grid(n, 8, times=times)
    st64 = device_grid(n, 8, lv64, times)
    lv32 = [(g.astype(np.float32), m.astype(np.float32)) for g, m in lv64]
    st32 = device_grid(n, 4, lv32, times)
    out = []
    for st, prec in ((st64, G.FP64), (st32, G.FP32)):
        with G.Render(24, 24, precision=prec, spacetime=st) as r:
            r.icond(32.0, 32.0, 35.0, 45.0, 0.0)
            r.render(-32.0, 1024, 3)
            out.append((r.read_f64()[0], r.steps()))
    (d64, s64), (d32, s32) = out
    same = s64 == s32
    assert same.mean() > 0.9
    dx = np.abs(d32[same, 1:4] - d64[same, 1:4]).max(axis=1)
    assert np.median(dx) < 1e-3 and dx.max() < 0.5, (np.median(dx), dx.max())


def test_launcher_dyst_job_matches_oracle(tmp_path):
    exe = os.path.join(os.path.dirname(J.LIB_PATH), "gray-mkgrid")
    f = str(tmp_path / "bks.h5")
    r = subprocess.run([exe, f"out={f}", "n=33", "a=0.7", "v=0.3", "axis=x", "dtype=4",
                        "times=0.0"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([J.LAUNCHER, "gray", "coor=dyst", f"data={f}", "W=8", "H=8", "w=32",
                        "h=32", "r=35", "i=45", "p=8", "N=2", f"out={tmp_path}/%04d.raw"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "freezing spacetime" in r.stdout
    xs, ys, zs, dt = S.h5_read_axes(f)
    g, m = S.h5_read_level(f, "0.0", (33, 33, 33))
    og = O.GridSpacetime((33,) * 3, (-40.0,) * 3, (40.0,) * 3, [g], [m])
    d0, i0, s0 = og.render(8, 8, 32.0, 32.0, 35.0, 45.0, 0.0, n_dump=2)
    states, diag = J.load_raw(f"{tmp_path}/0002.raw")
    assert np.array_equal(states.reshape(-1, 8), d0)
    assert np.array_equal(diag.ravel(), i0)


def test_spacetime_errors():
    with pytest.raises(G.GrayHipError):
        G.Spacetime((1, 4, 4), (-1,) * 3, (1,) * 3)
    st = G.Spacetime((4, 4, 4), (-1,) * 3, (1,) * 3)
    with pytest.raises(G.GrayHipError):
        st.set_frozen(False)  # only one level could be loaded
    with pytest.raises(G.GrayHipError):   # gridded transfer uses the reference flow
        G.Render(4, 4, nu=(1e11,), spacetime=st, fluid=G.FLUID_DISK)
    with pytest.raises(G.GrayHipError):   # the thermal flow needs a numerical spacetime
        G.Render(4, 4, nu=(1e11,), fluid=G.FLUID_THERMAL)
    with G.Render(4, 4, nu=(1e11,), spacetime=st) as r:
        with pytest.raises(G.GrayHipError):   # no density uploaded
            r.icond(1.0, 1.0, 0.5, 45.0, 0.0)
    st.close()


def test_launcher_dyst_sliding_window_matches_oracle(tmp_path):
    """Five levels at t = -80..0 and three intervals of -32 from t_init = 0:
    the window slides down by one level (shift + the prefetched level), then
    jumps (both levels reloaded), then clamps at the first level
    (sim-org/gray.c:259-293).  The oracle replays the same windows."""
    exe = os.path.join(os.path.dirname(J.LIB_PATH), "gray-mkgrid")
    f = str(tmp_path / "bks5.h5")
    r = subprocess.run([exe, f"out={f}", "n=33", "a=0.7", "v=0.3", "axis=x", "dtype=8",
                        "times=-80.0,-60.0,-40.0,-20.0,0.0"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([J.LAUNCHER, "gray", "coor=dyst", f"data={f}", "W=8", "H=8", "w=32",
                        "h=32", "r=35", "i=45", "p=8", "N=3", "t0=0",
                        f"out={tmp_path}/%04d.raw"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "slow light" in r.stdout
    names, times = S.h5_list_times(f)
    order = np.argsort(times)
    names, times = [names[k] for k in order], times[order]
    lv = [S.h5_read_level(f, nm, (33, 33, 33)) for nm in names]

    def find(t):
        if t <= times[0]:
            return 0
        for l in range(len(times) - 1):
            if times[l] <= t <= times[l + 1]:
                return l
        return len(times) - 2

    def window(s):
        return O.GridSpacetime((33,) * 3, (-40.0,) * 3, (40.0,) * 3, [lv[s][0], lv[s + 1][0]],
                               [lv[s][1], lv[s + 1][1]], times=(times[s], times[s + 1]))

    snap = find(0.0)
    data, info = window(snap).icond_all(8, 8, 32.0, 32.0, 35.0, 45.0, 0.0)
    steps = np.zeros(64, dtype=np.int64)
    seen = [snap]
    for i in range(1, 4):
        target = -32.0 * i
        G0 = window(snap)
        G0.evolve_all(data, info, -32.0, 1024, steps)
        if target < times[snap] or target > times[snap + 1]:
            snap = find(target)
            seen.append(snap)
        states, diag = J.load_raw(f"{tmp_path}/{i:04d}.raw")
        assert np.array_equal(states.reshape(-1, 8), data), i
        assert np.array_equal(diag.ravel(), info), i
    assert seen == [3, 2, 0, 0]


def test_fp32_render_matches_interval_launches():
    """The FP32 form's one launch over 3 intervals against 3 launches of one
    interval each.  Between launches the compensated state is stored as
    hi + lo in FP64, which rounds a lo more than 2^29 below hi, so the two
    agree to FP32-form accuracy, not bit for bit."""
    n, times = 33, (0.0, 1.0)
    xs, lv = make_grid(n, 4, times=times)
    st = device_grid(n, 4, lv, times)
    out = []
    for per_launch in (False, True):
        with G.Render(20, 12, precision=G.FP32, spacetime=st, t_init=1.0) as r:
            r.icond(32.0, 32.0, 35.0, 45.0, 0.0)
            if per_launch:
                for _ in range(3):
                    r.evolve(-32.0, 1024)
            else:
                r.render(-32.0, 1024, 3)
            out.append((r.read_f64()[0], r.steps()))
    st.close()
    (d0, s0), (d1, s1) = out
    same = s0 == s1
    assert same.mean() > 0.95
    dx = np.abs(d1[same, :4] - d0[same, :4]).max(axis=1)
    assert dx.max() < 1e-4, dx.max()


@pytest.mark.parametrize("precision", [G.FP32, G.FP64])
@pytest.mark.parametrize("dtype", [4, 8])
def test_bricks_do_not_change_results(dtype, precision, monkeypatch):
    """The tile kernels' LDS bricks hold the grid's own values and are summed
    in the global path's order: the render with bricks equals the render
    with every corner read from global memory, bit for bit (40x36 rays: 30
    tiles of 8x8, the last row of tiles half outside the image)."""
    n, times = 33, (0.0, 1.0)
    xs, lv = make_grid(n, dtype, times=times)
    st = device_grid(n, dtype, lv, times)
    out = []
    for flag in ("1", "0"):
        monkeypatch.setenv("GRAYHIP_GRID_BRICKS", flag)
        with G.Render(40, 36, precision=precision, spacetime=st, t_init=1.0) as r:
            r.icond(32.0, 32.0, 35.0, 45.0, 0.0)
            r.render(-32.0, 1024, 3)
            out.append((r.read_f64(), r.steps()))
    st.close()
    (d0, i0), s0 = out[0]
    (d1, i1), s1 = out[1]
    assert np.array_equal(s0, s1)
    assert np.array_equal(d0, d1)
    assert np.array_equal(i0, i1)


@pytest.mark.parametrize("dtype", [4, 8])
def test_covariant_view_bitwise(dtype):
    """k_mu = g(q) u with the interpolated metric (dyst.cl down) of the final
    states of a two-level render: the grid oracle's, bit for bit."""
    n, times = 33, (0.0, 1.0)
    xs, lv = make_grid(n, dtype, times=times)
    st = device_grid(n, dtype, lv, times)
    with G.Render(11, 9, precision=G.FP64, spacetime=st, t_init=1.0) as r:
        r.icond(32.0, 32.0, 35.0, 45.0, 0.0)
        r.render(-32.0, 1024, 2)
        d1, _ = r.read_f64()
        cov = r.read_covariant()
    st.close()
    og = oracle_grid(n, lv, times, t_off=1.0)
    want = np.array([np.concatenate([d1[g, :4], og.down(d1[g, :4], d1[g, 4:8])])
                     for g in range(len(d1))])
    assert np.array_equal(cov, want)
