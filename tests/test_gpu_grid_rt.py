"""Radiative transfer on numerical spacetimes with the reference's flow and
emission model (sim-org/flow.cl:36-58, rt.cl:80-198) on the GPU, through
the C ABI.

Bars:
  * plain FP64 form against the FP64 grid oracle: geodesic state and step
    counts bit for bit; I and tau to 1e-10 relative (exp and log are not
    correctly rounded on either side);
  * flat grid (Minkowski metric, zero Christoffels, constant density): the
    closed forms I = j_ff L and tau = j_ff L / B_nu at path length L, in both
    forms (FP64 1e-12, FP32 form 1e-5 relative);
  * FP32 form (FP32 grid, state scaled by K = M_ADM ff_f^2) against the FP64
    form: image relative L1 per channel <= 1e-4.
The same bars hold with a magnetic field b^mu = (0, 0, 0, B) in the flow
(flow.cl:47 has B = 0), which adds the synchrotron term of rt.cl:145-170;
on the flat grid the closed form becomes I = (j_ff + j_syn(cos)) L with
cos = k_z B / (shift B + eps) = u^z / u^t constant along each ray.
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
NUS = (1e11, 2.3e11, 5e11, 8e11)


def boosted(n, times, dtype, rho_cap=100.0):
    """Boosted Kerr-Schild levels with the generator's density 1/|x|, capped
    at rho_cap: uncapped, the grid point at the origin holds ~1.8e308 (the
    generator's nan_to_num of 1/0), and rays through its cells carry
    j ~ n_e^2 = inf."""
    xs = S.axis(-BOX, BOX, n)
    b = S.bks(1.0, 0.7, 0.3, (1, 0, 0))
    lv = [S.bks_sample(b, t, xs, xs, xs, rho=True) for t in times]
    if rho_cap:
        lv = [(g, m, np.minimum(r, rho_cap)) for g, m, r in lv]
    tp = np.float32 if dtype == 4 else np.float64
    return [tuple(a.astype(tp) for a in l) for l in lv]


def same_rt(rt0, rt1, rel=1e-10):
    """Equal non-finite pattern; finite values within rel."""
    fin0, fin1 = np.isfinite(rt0), np.isfinite(rt1)
    assert np.array_equal(fin0, fin1)
    assert np.array_equal(rt0[~fin0], rt1[~fin1], equal_nan=True)
    a, b = rt0[fin0], rt1[fin1]
    assert np.all(np.abs(b - a) <= rel * np.abs(a)), np.abs(b - a).max()


def device(n, lv, times, dtype, lo=-BOX, hi=BOX):
    st = G.Spacetime((n,) * 3, (lo,) * 3, (hi,) * 3, dtype=dtype)
    for slot, (g, m, r) in enumerate(lv):
        st.upload(slot, times[slot], g, m)
        st.upload_rho(slot, r)
    st.set_frozen(len(lv) == 1)
    return st


def gpu_render(st, precision, w, nus=NUS, t_init=0.0, n_dump=2, cam=(32.0, 35.0, 45.0),
               dt=-32.0, n_sub=1024, b_field=0.0):
    with G.Render(w, w, precision=precision, spacetime=st, nu=nus, t_init=t_init,
                  b_field=b_field) as r:
        r.icond(cam[0], cam[0], cam[1], cam[2], 0.0)
        for _ in range(n_dump):
            r.evolve(dt, n_sub)
        d, i = r.read_f64()
        return d, i, r.steps()


@pytest.mark.parametrize("dtype,times,cap", [(8, (0.0,), 100.0), (4, (0.0, 1.0), 100.0),
                                             (8, (0.0,), None)])
def test_thermal_fp64_matches_oracle(dtype, times, cap):
    n = 33
    lv = boosted(n, times, dtype, cap)
    st = device(n, lv, times, dtype)
    t_init = times[-1]
    d1, i1, s1 = gpu_render(st, G.FP64, 12, t_init=t_init)
    st.close()
    og = O.GridSpacetime((n,) * 3, (-BOX,) * 3, (BOX,) * 3,
                         [l[0].astype(np.float64) for l in lv], [l[1].astype(np.float64) for l in lv],
                         times=(times[0], times[-1]), t_off=t_init,
                         rho=[l[2].astype(np.float64) for l in lv], nu=NUS)
    d0, i0, s0 = og.render(12, 12, 32.0, 32.0, 35.0, 45.0, 0.0, n_dump=2)
    assert np.array_equal(s0, s1)
    assert np.array_equal(d0[:, :8], d1[:, :8])
    assert np.array_equal(i0, i1)
    rt0, rt1 = d0[:, 8:], d1[:, 8:]
    assert (np.abs(rt0) > 0).any()
    if cap:
        assert np.isfinite(rt0).all()
    same_rt(rt0, rt1)


@pytest.mark.parametrize("precision,tol", [(G.FP64, 1e-12), (G.FP32, 1e-5)])
def test_flat_slab_closed_form(precision, tol):
    n, L = 9, 10.0
    xs = np.linspace(-L, L, n)
    npts = n ** 3
    gam = np.zeros((npts, 40))
    met = np.zeros((npts, 10))
    met[:, 0], met[:, 4], met[:, 7], met[:, 9] = -1.0, 1.0, 1.0, 1.0
    rho = np.full(npts, 0.6)
    dtype = 8 if precision == G.FP64 else 4
    tp = np.float64 if dtype == 8 else np.float32
    st = device(n, [(gam.astype(tp), met.astype(tp), rho.astype(tp))], (0.0,), dtype, -L, L)
    d, _, s = gpu_render(st, precision, 8, cam=(4.0, 5.0, 45.0), dt=-4.0, n_sub=64, n_dump=6)
    st.close()
    rt = O.Rt()
    path = -d[:, 0]                          # u^t = 1: affine length = -q^t
    assert (path > 4.0).all()
    for i, v in enumerate(NUS):
        j = rt.j_ff(v, 1e12, 0.6)
        B = rt.b_planck(v, 1e12)
        np.testing.assert_allclose(d[:, 8 + i], j * path, rtol=tol)
        np.testing.assert_allclose(d[:, 8 + len(NUS) + i], j / B * path, rtol=tol)


def test_thermal_fp32_form_vs_fp64():
    n, times = 33, (0.0,)
    lv64 = boosted(n, times, 8)
    lv32 = [tuple(a.astype(np.float32) for a in l) for l in lv64]
    st64, st32 = device(n, lv64, times, 8), device(n, lv32, times, 4)
    d64, _, s64 = gpu_render(st64, G.FP64, 32, n_dump=3)
    d32, _, s32 = gpu_render(st32, G.FP32, 32, n_dump=3)
    st64.close()
    st32.close()
    nf = len(NUS)
    for blk in (slice(8, 8 + nf), slice(8 + nf, 8 + 2 * nf)):
        a, b = d32[:, blk], d64[:, blk]
        l1 = np.abs(a - b).sum(axis=0) / np.abs(b).sum(axis=0)
        assert (l1 <= 1e-4).all(), l1


def test_thermal_fp32_bricks_do_not_change_results(monkeypatch):
    """Transfer rides along the FP32 tile kernel's LDS-staged geodesics: the
    render with bricks equals the one with every corner read from global
    memory, bit for bit, channels included."""
    n, times = 33, (0.0, 1.0)
    st = device(n, boosted(n, times, 4), times, 4)
    out = []
    for flag in ("1", "0"):
        monkeypatch.setenv("GRAYHIP_GRID_BRICKS", flag)
        out.append(gpu_render(st, G.FP32, 32, t_init=1.0, n_dump=3))
    st.close()
    (d0, i0, s0), (d1, i1, s1) = out
    assert np.array_equal(s0, s1)
    assert np.array_equal(d0, d1, equal_nan=True)


def test_launcher_dyst_job_with_transfer(tmp_path):
    exe = os.path.join(os.path.dirname(J.LIB_PATH), "gray-mkgrid")
    f = str(tmp_path / "bks.h5")
    r = subprocess.run([exe, f"out={f}", "n=33", "a=0.7", "v=0.3", "axis=x", "dtype=8",
                        "times=0.0"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([J.LAUNCHER, "gray", "coor=dyst", f"data={f}", "W=8", "H=8", "w=32",
                        "h=32", "r=35", "i=45", "p=8", "N=2", "nu=2.3e11,5e11",
                        f"out={tmp_path}/%04d.raw"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    g, m, rho = S.h5_read_level(f, "0.0", (33, 33, 33), rho=True)
    og = O.GridSpacetime((33,) * 3, (-40.0,) * 3, (40.0,) * 3, [g], [m], rho=[rho],
                         nu=(2.3e11, 5e11))
    d0, i0, s0 = og.render(8, 8, 32.0, 32.0, 35.0, 45.0, 0.0, n_dump=2)
    states, diag = J.load_raw(f"{tmp_path}/0002.raw")
    states = states.reshape(-1, 12)
    assert np.array_equal(states[:, :8], d0[:, :8])
    same_rt(d0[:, 8:], states[:, 8:])


B_FIELD = 20.0


def test_synchrotron_fp64_matches_oracle():
    n, times = 33, (0.0, 1.0)
    lv = boosted(n, times, 8)
    st = device(n, lv, times, 8)
    d1, i1, s1 = gpu_render(st, G.FP64, 12, t_init=1.0, b_field=B_FIELD)
    d2, _, _ = gpu_render(st, G.FP64, 12, t_init=1.0)
    st.close()
    og = O.GridSpacetime((n,) * 3, (-BOX,) * 3, (BOX,) * 3, [l[0] for l in lv], [l[1] for l in lv],
                         times=times, t_off=1.0, rho=[l[2] for l in lv], nu=NUS, b_field=B_FIELD)
    d0, i0, s0 = og.render(12, 12, 32.0, 32.0, 35.0, 45.0, 0.0, n_dump=2)
    assert np.array_equal(s0, s1)
    assert np.array_equal(d0[:, :8], d1[:, :8])
    assert np.array_equal(i0, i1)
    same_rt(d0[:, 8:], d1[:, 8:])
    # the field's emission dominates the image: without it the intensities
    # are those of bremsstrahlung alone, many orders of magnitude fainter
    nf = len(NUS)
    assert d1[:, 8:8 + nf].sum() > 1e6 * d2[:, 8:8 + nf].sum() > 0


@pytest.mark.parametrize("precision,tol", [(G.FP64, 1e-12), (G.FP32, 1e-5)])
def test_flat_slab_with_field_closed_form(precision, tol):
    n, L = 9, 10.0
    npts = n ** 3
    gam = np.zeros((npts, 40))
    met = np.zeros((npts, 10))
    met[:, 0], met[:, 4], met[:, 7], met[:, 9] = -1.0, 1.0, 1.0, 1.0
    rho = np.full(npts, 0.6)
    dtype = 8 if precision == G.FP64 else 4
    tp = np.float64 if dtype == 8 else np.float32
    st = device(n, [(gam.astype(tp), met.astype(tp), rho.astype(tp))], (0.0,), dtype, -L, L)
    # a camera's rays are parallel: two inclinations give two field angles
    d = np.concatenate([gpu_render(st, precision, 8, cam=(4.0, 5.0, inc), dt=-4.0, n_sub=64,
                                   n_dump=6, b_field=B_FIELD)[0] for inc in (45.0, 75.0)])
    st.close()
    rt = O.Rt()
    path = -d[:, 0]
    cos = d[:, 7] / d[:, 4]                  # k_z / -k_t in flat space
    assert (path > 4.0).all() and (np.abs(cos) < 0.99).all() and np.ptp(cos) > 0.4
    for i, v in enumerate(NUS):
        js = np.array([rt.j_syn(v, 1e12, 0.6, B_FIELD, c) for c in cos])
        j = rt.j_ff(v, 1e12, 0.6) + js
        B = rt.b_planck(v, 1e12)
        assert (js > 1e10 * rt.j_ff(v, 1e12, 0.6)).all()
        np.testing.assert_allclose(d[:, 8 + i], j * path, rtol=tol)
        np.testing.assert_allclose(d[:, 8 + len(NUS) + i], j / B * path, rtol=tol)


def test_synchrotron_fp32_form_vs_fp64():
    n, times = 33, (0.0,)
    lv64 = boosted(n, times, 8)
    lv32 = [tuple(a.astype(np.float32) for a in l) for l in lv64]
    st64, st32 = device(n, lv64, times, 8), device(n, lv32, times, 4)
    d64, _, _ = gpu_render(st64, G.FP64, 32, n_dump=3, b_field=B_FIELD)
    d32, _, _ = gpu_render(st32, G.FP32, 32, n_dump=3, b_field=B_FIELD)
    st64.close()
    st32.close()
    nf = len(NUS)
    for blk in (slice(8, 8 + nf), slice(8 + nf, 8 + 2 * nf)):
        a, b = d32[:, blk], d64[:, blk]
        l1 = np.abs(a - b).sum(axis=0) / np.abs(b).sum(axis=0)
        assert (l1 <= 1e-4).all(), l1


def test_field_needs_the_thermal_flow():
    with pytest.raises(G.GrayHipError):
        G.Render(4, 4, 0.5, precision=G.FP64, nu=(1.0,), b_field=1.0).close()
    with pytest.raises(G.GrayHipError):
        G.Render(4, 4, 0.5, precision=G.FP64, b_field=-1.0).close()
