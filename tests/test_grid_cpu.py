"""Numerical-spacetime path on the CPU: the boosted Kerr-Schild generator,
the grid oracle (interpolation, RHS, camera, step rule) and HDF5 files.

Pins of the grid oracle (the reference holds no fixtures for this path):
  * trilinear-in-space interpolation is exact on multilinear fields, on
    uniform and fisheye axes; clamp-to-edge outside the box;
  * linear-in-time blending between the two levels, clamped, and
    space-only when the levels coincide;
  * on a grid sampled from the unboosted Kerr-Schild metric, the gridded RHS
    converges to the analytic RHS at second order in the spacing;
  * the camera direction is null for the interpolated metric;
and of the generator: at v = 0 its metric and Christoffels reproduce the
analytic oracle (down, gr_rhs); boosted, its metric transforms as a tensor
under the boost and its Christoffels match finite differences.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from gray_amd import spacetime as S
from oracle import oracle as O

IDX = [[0, 1, 2, 3], [1, 4, 5, 6], [2, 5, 7, 8], [3, 6, 8, 9]]


def full(met10):
    return np.array([[met10[IDX[a][b]] for b in range(4)] for a in range(4)])


def gamma_full(gam40):
    return np.array([[[gam40[10 * m + IDX[a][b]] for b in range(4)] for a in range(4)]
                     for m in range(4)])


@pytest.mark.parametrize("a", [0.0, 0.7, 0.9375])
def test_generator_unboosted_matches_analytic_oracle(a):
    b = S.bks(1.0, a, 0.0)
    rng = np.random.default_rng(11)
    for _ in range(20):
        q = np.concatenate([[rng.uniform(-5, 5)], rng.uniform(-12, 12, 3)])
        if np.linalg.norm(q[1:]) < 2.5:
            continue
        u = rng.normal(size=4)
        met, gam = S.bks_point(b, q)
        np.testing.assert_allclose(full(met) @ u, O.down(a, q, u), rtol=1e-13, atol=1e-14)
        du = -np.einsum("mab,a,b->m", gamma_full(gam), u, u)
        ref = O.gr_rhs(a, q, u)[1]
        np.testing.assert_allclose(du, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("axis", [(1, 0, 0), (0, 0, 1), (0.6, 0.8, 0)])
def test_generator_boost_covariance(axis):
    a, v = 0.7, 0.3
    n = np.array(axis, float) / np.linalg.norm(axis)
    gam_ = 1 / math.sqrt(1 - v * v)
    L = np.eye(4)  # rest = L @ lab
    L[0, 0] = gam_
    L[0, 1:] = -gam_ * v * n
    L[1:, 0] = -gam_ * v * n
    L[1:, 1:] = np.eye(3) + (gam_ - 1) * np.outer(n, n)
    rest, lab = S.bks(1.0, a, 0.0), S.bks(1.0, a, v, axis)
    for x in ([0.3, 4.0, 1.0, -2.0], [1.0, -3.0, 5.0, 2.0], [-2.0, 7.0, -7.0, 1.0]):
        x = np.array(x)
        g_lab = full(S.bks_point(lab, x)[0])
        g_rest = full(S.bks_point(rest, L @ x)[0])
        np.testing.assert_allclose(g_lab, L.T @ g_rest @ L, rtol=1e-12, atol=1e-13)
        assert np.linalg.det(g_lab) == pytest.approx(-1.0, rel=1e-12)


def test_generator_christoffels_match_finite_differences():
    b = S.bks(1.0, 0.7, 0.3, (1, 0, 0))
    x = np.array([0.5, 3.0, -2.0, 1.5])
    h = 1e-4
    dg = []
    for c in range(4):
        e = np.zeros(4)
        e[c] = h
        gp2, gp1 = full(S.bks_point(b, x + 2 * e)[0]), full(S.bks_point(b, x + e)[0])
        gm1, gm2 = full(S.bks_point(b, x - e)[0]), full(S.bks_point(b, x - 2 * e)[0])
        dg.append((-gp2 + 8 * gp1 - 8 * gm1 + gm2) / (12 * h))
    met, gam = S.bks_point(b, x)
    gi = np.linalg.inv(full(met))
    G = gamma_full(gam)
    for m in range(4):
        for al in range(4):
            for be in range(4):
                ref = 0.5 * sum(gi[m, s] * (dg[be][s, al] + dg[al][s, be] - dg[s][al, be])
                                for s in range(4))
                assert G[m, al, be] == pytest.approx(ref, rel=1e-7, abs=1e-9)


def multilinear(x, y, z, c):
    return (1 + 2 * x - 3 * y + 0.5 * z + 0.1 * x * y - 0.2 * x * z + 0.05 * y * z
            + 0.01 * x * y * z) * (1 + 0.1 * c)


def grid_of(xs, ys, zs, fn_g, fn_m):
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    X, Y, Z = (A.transpose(2, 1, 0).ravel() for A in (X, Y, Z))  # x fastest
    gam = np.stack([fn_g(X, Y, Z, c) for c in range(40)], axis=1)
    met = np.stack([fn_m(X, Y, Z, c) for c in range(10)], axis=1)
    return gam, met


@pytest.mark.parametrize("fisheye", [False, True])
def test_trilinear_exact_on_multilinear_fields(fisheye):
    n = (9, 7, 8)
    if fisheye:
        xs, ys, zs = S.axis(-30, 40, n[0], True), S.axis(-20, 25, n[1], True), S.axis(-35, 30, n[2], True)
        assert S.axis_map(xs) == 1
    else:
        xs, ys, zs = S.axis(-4, 5, n[0]), S.axis(-3, 2, n[1]), S.axis(-6, 1, n[2])
        assert S.axis_map(xs) == 0
    gam, met = grid_of(xs, ys, zs, multilinear, lambda X, Y, Z, c: multilinear(Y, Z, X, c))
    G = O.GridSpacetime(n, (xs[0], ys[0], zs[0]), (xs[-1], ys[-1], zs[-1]), [gam], [met],
                        map=O.MAP_FISHEYE if fisheye else O.MAP_UNIFORM)
    rng = np.random.default_rng(5)
    for _ in range(50):
        q = np.array([0.0, rng.uniform(xs[0], xs[-1]), rng.uniform(ys[0], ys[-1]),
                      rng.uniform(zs[0], zs[-1])])
        for c in (0, 17, 39):
            want = multilinear(q[1], q[2], q[3], c)
            assert G.sample(0, c, q) == pytest.approx(want, rel=1e-11, abs=1e-11)
        assert G.sample(1, 3, q) == pytest.approx(multilinear(q[2], q[3], q[1], 3), rel=1e-11)
    # clamp to edge outside the box
    q = np.array([0.0, xs[-1] + 3.0, ys[0] - 2.0, zs[0] + 0.0])
    assert G.sample(0, 0, q) == pytest.approx(multilinear(xs[-1], ys[0], zs[0], 0), rel=1e-12)


def test_time_blend_and_clamp():
    n = (4, 4, 4)
    xs = S.axis(-1, 1, 4)
    g1, m1 = grid_of(xs, xs, xs, lambda X, Y, Z, c: 0 * X + 1.0, lambda X, Y, Z, c: 0 * X + 1.0)
    g2, m2 = g1 + 2.0, m1 + 2.0
    G = O.GridSpacetime(n, (-1,) * 3, (1,) * 3, [g1, g2], [m1, m2], times=(10.0, 12.0), t_off=10.0)
    for qt, want in [(0.0, 1.0), (1.0, 2.0), (0.5, 1.5), (-5.0, 1.0), (7.0, 3.0)]:
        assert G.sample(0, 5, [qt, 0.1, 0.2, -0.3]) == pytest.approx(want, rel=1e-15)
    Gf = O.GridSpacetime(n, (-1,) * 3, (1,) * 3, [g1, g2], [m1, m2], times=(10.0, 10.0))
    assert Gf.sample(0, 5, [1.0, 0.1, 0.2, -0.3]) == pytest.approx(1.0, rel=1e-15)


@pytest.mark.parametrize("a", [0.0, 0.9])
def test_grid_rhs_converges_to_analytic_second_order(a):
    b = S.bks(1.0, a, 0.0)
    q = np.array([0.0, 4.3, -2.1, 3.7])
    u = np.array([1.0, 0.4, 0.3, -0.5])
    ref = O.gr_rhs(a, q, u)[1]
    ref = ref - ref[0] * u  # coordinate-time form of dyst.cl:186-191
    errs = []
    for h in (0.4, 0.2, 0.1):
        axes = [q[i + 1] + h * (np.arange(8) - 3.5) for i in range(3)]
        gam, met = S.bks_sample(b, 0.0, *axes)
        G = O.GridSpacetime((8, 8, 8), [ax[0] for ax in axes], [ax[-1] for ax in axes],
                            [gam], [met])
        errs.append(np.abs(G.rhs(q, u) - ref).max())
    assert errs[0] / errs[1] > 3.0 and errs[1] / errs[2] > 3.0, errs
    assert errs[2] < 1e-3 * np.abs(ref).max()


def test_grid_icond_null_step_rule_and_box_exit():
    b = S.bks(1.0, 0.7, 0.3, (1, 0, 0))
    xs = S.axis(-40, 40, 33)
    gam, met = S.bks_sample(b, 0.0, xs, xs, xs)
    G = O.GridSpacetime((33,) * 3, (-40,) * 3, (40,) * 3, [gam], [met])
    q, u = G.icond(35.0, 45.0, 0.0, 3.0, -5.0)
    assert u[0] == 1.0
    assert abs(G.getuu(q, u)) < 1e-13
    Graw = O.GridSpacetime((33,) * 3, (-40,) * 3, (40,) * 3, [gam], [met], raw_icond=True)
    q2, u2 = Graw.icond(35.0, 45.0, 0.0, 3.0, -5.0)
    assert np.array_equal(q, q2) and np.linalg.norm(u2[1:]) == pytest.approx(1.0)
    assert G.getdt(q, -0.03125) == -0.01
    assert G.getdt(q, -0.005) == -0.005
    assert G.getdt([0.0, 40.5, 0.0, 0.0], -0.03125) == 0.0
    Gref = O.GridSpacetime((33,) * 3, (-40,) * 3, (40,) * 3, [gam], [met], ref_rule=True)
    assert Gref.getdt([0.0, 40.5, 0.0, 0.0], -0.03125) == -0.01
    # horizon proxy: coordinate speed |dx/dt| above 4 stops the ray; NaN too
    assert G.getdt(q, -0.03125, u=[1.0, 0.0, 4.0, 0.0]) == -0.01
    assert G.getdt(q, -0.03125, u=[1.0, 0.0, 4.0, 1e-7]) == 0.0
    assert G.getdt(q, -0.03125, u=[1.0, np.nan, 0.0, 0.0]) == 0.0
    assert Gref.getdt(q, -0.03125, u=[1.0, 0.0, 40.0, 0.0]) == -0.01


def test_grid_horizon_proxy_stops_only_captured_rays():
    """Ten pixels of the 512^2, 32 M x 32 M, r_obs = 35, i = 45 image of the
    128^3 boosted Kerr-Schild grid (a = 0.7, v = 0.3 along x, levels 0 and
    1) whose rays wind into the hole.  Under the reference rule they run
    into the singularity and overflow to NaN; under the default rule the
    speed cap stops each of them early, finite, inside the horizon
    (|x| < 1.85 for a = 0.7)."""
    n = 128
    xs = S.axis(-40, 40, n)
    b = S.bks(1.0, 0.7, 0.3, (1, 0, 0))
    lv = [S.bks_sample(b, t, xs, xs, xs) for t in (0.0, 1.0)]
    pix = [(288, 220), (292, 220), (220, 224), (292, 224), (296, 224), (220, 228),
           (292, 228), (296, 228), (284, 232), (288, 232)]
    out = []
    for ref in (False, True):
        G = O.GridSpacetime((n,) * 3, (-40.0,) * 3, (40.0,) * 3, [lv[0][0], lv[1][0]],
                            [lv[0][1], lv[1][1]], times=(0.0, 1.0), t_off=1.0, ref_rule=ref)
        rows = []
        for gi, gj in pix:
            q, u = G.icond(35.0, 45.0, 0.0, ((gi + 0.5) / 512 - 0.5) * 32.0,
                           ((gj + 0.5) / 512 - 0.5) * 32.0)
            rows.append(np.concatenate([q, u]))
        d = np.array(rows)
        steps = np.zeros(len(d), dtype=np.int64)
        for _ in range(4):
            G.evolve_all(d, np.zeros(len(d)), -32.0, 1024, steps, 8)
        out.append((d, steps))
    (d, steps), (dref, _) = out
    assert not np.isfinite(dref).all(axis=1).any()
    assert np.isfinite(d).all()
    assert (steps < 4 * 3201).all()
    assert (np.linalg.norm(d[:, 1:4], axis=1) < 1.85).all()


def test_grid_ray_leaves_the_box():
    b = S.bks(1.0, 0.7, 0.3, (1, 0, 0))
    xs = S.axis(-40, 40, 41)
    gam, met = S.bks_sample(b, 0.0, xs, xs, xs)
    G = O.GridSpacetime((41,) * 3, (-40,) * 3, (40,) * 3, [gam], [met])
    data, info, steps = G.render(3, 3, 32.0, 32.0, 35.0, 45.0, 0.0, n_dump=3)
    far = np.abs(data[:, 1:4]).max(axis=1)
    # 3 intervals of 3200 steps of 0.01 plus a 1024-unit remainder step each
    assert (steps <= 3 * 3201).all()
    stopped = steps < 3 * 3201
    assert stopped.sum() >= 3 and (far[stopped] > 40.0).all()  # left the box
    assert (far[~stopped] < 40.0).all()                        # still inside


def test_h5_round_trip(tmp_path):
    b = S.bks(1.0, 0.7, 0.3, (1, 0, 0))
    n = (6, 5, 4)
    xs, ys, zs = S.axis(-3, 3, 6), S.axis(-2, 2, 5), S.axis(-1.5, 1.5, 4)
    for dtype in (4, 8):
        p = str(tmp_path / f"s{dtype}.h5")
        S.h5_create(p, xs, ys, zs, dtype=dtype)
        lv = {}
        for t in ("10.0", "2.0", "1.0"):  # written out of order
            gam, met, rho = S.bks_sample(b, float(t), xs, ys, zs, rho=True)
            S.h5_write_level(p, t, n, gam, met, rho, dtype=dtype)
            lv[t] = (gam, met, rho)
        names, times = S.h5_list_times(p)
        assert names == ["1.0", "2.0", "10.0"] and list(times) == [1.0, 2.0, 10.0]
        rx, ry, rz, dt = S.h5_read_axes(p)
        assert dt == dtype
        np.testing.assert_array_equal(rx, xs.astype(np.float32 if dtype == 4 else np.float64))
        g, m, r = S.h5_read_level(p, "2.0", n, dtype=8, rho=True)
        cast = (lambda a: a.astype(np.float32).astype(np.float64)) if dtype == 4 else (lambda a: a)
        np.testing.assert_array_equal(g, cast(lv["2.0"][0]))
        np.testing.assert_array_equal(m, cast(lv["2.0"][1]))
        np.testing.assert_array_equal(r, cast(lv["2.0"][2]))
        g4, m4 = S.h5_read_level(p, "1.0", n, dtype=4)
        assert g4.dtype == np.float32


def test_h5_errors(tmp_path):
    with pytest.raises(S.H5Error):
        S.h5_list_times(str(tmp_path / "missing.h5"))
    xs = S.axis(-1, 1, 3)
    p = str(tmp_path / "only_grid.h5")
    S.h5_create(p, xs, xs, xs)
    with pytest.raises(S.H5Error, match="not enough groups"):
        S.h5_list_times(p)
    b = S.bks(1.0, 0.5, 0.0)
    gam, met = S.bks_sample(b, 0.0, xs, xs, xs)
    S.h5_write_level(p, "0.0", (3, 3, 3), gam, met)
    with pytest.raises(S.H5Error, match="points"):
        S.h5_read_level(p, "0.0", (3, 3, 4))
    with pytest.raises(S.H5Error, match="rho"):  # rho was not written
        S.h5_read_level(p, "0.0", (3, 3, 3))


def test_mkgrid_cli(tmp_path):
    exe = os.path.join(os.path.dirname(S._job.LIB_PATH), "gray-mkgrid")
    p = str(tmp_path / "g.h5")
    r = subprocess.run([exe, f"out={p}", "n=8", "a=0.7", "v=0.3", "axis=x", "dtype=8"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    names, times = S.h5_list_times(p)
    assert names == ["0.0", "1.0"]
    xs, ys, zs, dt = S.h5_read_axes(p)
    g, m = S.h5_read_level(p, "1.0", (8, 8, 8))
    gam, met = S.bks_sample(S.bks(1.0, 0.7, 0.3, (1, 0, 0)), 1.0, xs, ys, zs)
    np.testing.assert_array_equal(g, gam)
