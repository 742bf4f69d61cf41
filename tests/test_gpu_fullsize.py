"""Parity at the full benchmark sizes, on strided subsets the oracle finishes
in seconds, plus the size-independent image properties of the whole frame.

  * analytic Kerr, 1024^2 rays, disk transfer with 4 channels (the headline
    workload): the plain FP64 form against the oracle on every 16th row and
    column, 4096 rays (geodesic state bit for bit, I and tau to 1e-10
    relative); the
    FP32 form against the plain FP64 form on the whole image (relative L1
    per channel <= 1e-4, peak-pixel error <= 5e-2, null-constraint drift of
    escaping rays <= 1e-3 and <= 1e-5);
  * 128^3 boosted Kerr-Schild grid with two time levels, 512^2 rays: the
    plain FP64 form against the grid oracle on every 16th row and column,
    1024 rays (state, step counts and g(u,u) bit for bit).
"""
import numpy as np
import pytest

from gray_amd import grayhip as G
from gray_amd import spacetime as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

A = 0.9375
NUS = (1.0, 2.67, 4.0, 8.0)


def kerr_render(precision, w=1024):
    with G.Render(w, w, A, precision=precision, nu=NUS) as r:
        r.icond(64.0, 64.0, 1000.0, 60.0, 0.0)
        r.render(-32.0, 1024, 128)
        d, info = r.read_f64()
        return d, info


def test_kerr_1024_fp64_subset_matches_oracle_and_fp32_image():
    d64, i64 = kerr_render(G.FP64)
    p = O.make_params(A, nu=NUS)
    data, _ = O.icond_all(p, 1024, 1024, 64.0, 64.0, 1000.0, 60.0, 0.0)
    sub = np.ascontiguousarray(data.reshape(1024, 1024, -1)[::16, ::16].reshape(4096, -1))
    O.evolve_subset(p, sub, 1, -32.0, 1024, 128)
    got = d64.reshape(1024, 1024, -1)[::16, ::16].reshape(4096, -1)
    assert np.array_equal(got[:, :8], sub[:, :8])
    rt0, rt1 = sub[:, 8:], got[:, 8:]
    assert (rt0[:, :4] > 0).sum() > 300
    assert np.all(np.abs(rt1 - rt0) <= 1e-10 * np.abs(rt0))

    d32, i32 = kerr_render(G.FP32)
    I64, I32 = d64[:, 8:12], d32[:, 8:12]
    l1 = np.abs(I32 - I64).sum(0) / np.abs(I64).sum(0)
    peak = np.abs(I32 - I64).max() / np.abs(I64).max()
    assert (l1 <= 1e-4).all(), l1
    assert peak <= 5e-2, peak
    esc = np.sqrt(np.sum(d64[:, 1:4] ** 2, axis=1)) > 999.0
    assert esc.sum() > 500000
    assert np.abs(i32[esc]).max() <= 1e-3 and np.abs(i64[esc]).max() <= 1e-5


def test_grid_128_fp64_subset_matches_oracle():
    n, times = 128, (0.0, 1.0)
    xs = S.axis(-40.0, 40.0, n)
    b = S.bks(1.0, 0.7, 0.3, (1, 0, 0))
    lv = [S.bks_sample(b, t, xs, xs, xs) for t in times]
    st = G.Spacetime((n,) * 3, (-40.0,) * 3, (40.0,) * 3, dtype=8)
    for slot, (g, m) in enumerate(lv):
        st.upload(slot, times[slot], g, m)
    st.set_frozen(False)
    with G.Render(512, 512, precision=G.FP64, spacetime=st, t_init=1.0) as r:
        r.icond(32.0, 32.0, 35.0, 45.0, 0.0)
        r.render(-32.0, 1024, 4)
        d1, i1 = r.read_f64()
        s1 = r.steps()
    st.close()
    og = O.GridSpacetime((n,) * 3, (-40.0,) * 3, (40.0,) * 3,
                         [g for g, _ in lv], [m for _, m in lv],
                         times=times, t_off=1.0)
    data, info = og.icond_all(512, 512, 32.0, 32.0, 35.0, 45.0, 0.0)
    pick = (np.arange(512)[::16, None] * 512 + np.arange(512)[None, ::16]).ravel()
    d0, i0 = np.ascontiguousarray(data[pick]), np.ascontiguousarray(info[pick])
    s0 = np.zeros(len(pick), dtype=np.int64)
    for _ in range(4):
        og.evolve_all(d0, i0, -32.0, 1024, s0)
    assert np.array_equal(s0, s1[pick])
    assert np.array_equal(d0, d1[pick])
    assert np.array_equal(i0, i1[pick])
