"""Pins of the FP64 oracle against closed forms, invariants and the
mathematics of the Kerr metric (the reference ships no golden vectors for
this path, so these are the pins).

Each check is chosen so that a plausible slip in the restatement of
sim-org/KS.cl, phys.cl, RK4.cl or driver.cl fails at least one of them:
  * getrr / metric components vs. Boyer-Lindquist closed forms
    (g_tt, g_t_phi, g_phi_phi fix f, l and the sign of the spin);
  * det g = -1 and l null (Kerr-Schild structure);
  * gr_rhs vs. -Gamma^mu_ab u^a u^b from finite-difference Christoffels of
    the metric (any dropped or mis-signed term of the regrouped RHS);
  * null initial data, camera geometry;
  * conservation of E, L_z and g(u,u) along RK4, 4th-order convergence;
  * integer step quantisation known answers;
  * Schwarzschild critical impact parameter 3*sqrt(3);
  * transfer integrator vs. the uniform-slab closed form.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def ks_point(a, r, th, ph):
    """Cartesian Kerr-Schild point of spheroidal (r, theta, phi)."""
    R = math.sqrt(r * r + a * a) * math.sin(th)
    return np.array([0.0, R * math.cos(ph), R * math.sin(ph), r * math.cos(th)])


def metric(a, q):
    """g_{mu nu} assembled column by column from the oracle's down()."""
    return np.array([O.down(a, q, e) for e in np.eye(4)]).T


@pytest.mark.parametrize("a", [0.0, 0.5, 0.9375, 0.998])
def test_getrr_closed_form(a):
    for r, th, ph in [(1.5, 0.3, 0.1), (3.0, 1.2, 2.0), (10.0, 2.5, -1.0), (1000.0, 1.0472, 0.0)]:
        q = ks_point(a, r, th, ph)
        assert O.getrr(a, q) == pytest.approx(r * r, rel=1e-13)
        assert O.geteps(a, q) == pytest.approx(r - (1 + math.sqrt(1 - a * a)), rel=1e-12, abs=1e-13)


def test_horizon_and_isco_values():
    assert O.make_params(0.9375).r_h == pytest.approx(1.347985, abs=1e-6)
    assert O.make_params(0.998).r_h == pytest.approx(1.063214, abs=1e-6)
    assert O.lib().og_r_isco(0.0) == pytest.approx(6.0, rel=1e-14)
    assert O.lib().og_r_isco(1.0) == pytest.approx(1.0, rel=1e-12)
    assert O.lib().og_r_isco(0.9375) == pytest.approx(2.044201, abs=1e-6)
    assert O.lib().og_r_isco(0.998) == pytest.approx(1.236971, abs=1e-6)


@pytest.mark.parametrize("a", [0.0, 0.6, 0.9375])
def test_metric_boyer_lindquist_components(a):
    """g_tt, g_t phi and g_phi phi are the same in Kerr-Schild and
    Boyer-Lindquist coordinates; with d/dphi = (0, -y, x, 0)."""
    for r, th, ph in [(2.2, 0.7, 0.4), (5.0, 1.5707963, 1.0), (17.0, 2.8, -2.0)]:
        q = ks_point(a, r, th, ph)
        g = metric(a, q)
        S = r * r + a * a * math.cos(th) ** 2
        s2 = math.sin(th) ** 2
        xi = np.array([0.0, -q[2], q[1], 0.0])
        et = np.array([1.0, 0, 0, 0])
        assert et @ g @ et == pytest.approx(-(1 - 2 * r / S), rel=1e-12)
        assert et @ g @ xi == pytest.approx(-2 * a * r * s2 / S, rel=1e-11, abs=1e-14)
        assert xi @ g @ xi == pytest.approx((r * r + a * a + 2 * a * a * r * s2 / S) * s2, rel=1e-12)


@pytest.mark.parametrize("a", [0.3, 0.9375])
def test_kerr_schild_structure(a):
    eta = np.diag([-1.0, 1, 1, 1])
    for r, th, ph in [(1.8, 0.5, 0.3), (4.0, 2.0, 2.5), (60.0, 1.0, -0.7)]:
        q = ks_point(a, r, th, ph)
        g = metric(a, q)
        # (f lx) ly and (f ly) lx round differently: symmetric to an ulp
        assert np.allclose(g, g.T, rtol=4e-16, atol=1e-300)
        assert np.linalg.det(g) == pytest.approx(-1.0, rel=1e-12)
        # l^mu = eta^{mu nu} l_nu with l_t = 1 is null for g and for eta
        h = g - eta
        f = h[0, 0]
        l = h[0] / f
        lu = eta @ l
        assert l @ lu == pytest.approx(0.0, abs=1e-13)
        assert f == pytest.approx(2 * r ** 3 / (r ** 4 + a * a * q[3] ** 2), rel=1e-13)


def christoffel_fd(a, q, h=1e-4):
    """Gamma^mu_{ab} from 4th-order central differences of the metric."""
    dg = np.zeros((4, 4, 4))  # dg[c] = d g / d x^c
    for c in range(1, 4):
        e = np.zeros(4)
        e[c] = h
        gp2, gp1 = metric(a, q + 2 * e), metric(a, q + e)
        gm1, gm2 = metric(a, q - e), metric(a, q - 2 * e)
        dg[c] = (-gp2 + 8 * gp1 - 8 * gm1 + gm2) / (12 * h)
    gi = np.linalg.inv(metric(a, q))
    G = np.zeros((4, 4, 4))
    for m in range(4):
        for al in range(4):
            for be in range(4):
                G[m, al, be] = 0.5 * sum(gi[m, s] * (dg[be][s, al] + dg[al][s, be] - dg[s][al, be])
                                         for s in range(4))
    return G


@pytest.mark.parametrize("a", [0.0, 0.5, 0.9375])
def test_gr_rhs_matches_christoffel_contraction(a):
    rng = np.random.default_rng(7)
    for r, th, ph in [(2.5, 0.9, 0.2), (4.0, 1.5, 2.0), (9.0, 2.4, -1.3)]:
        q = ks_point(a, r, th, ph)
        u = rng.normal(size=4)
        dq, du = O.gr_rhs(a, q, u)
        G = christoffel_fd(a, q)
        ref = -np.einsum("mab,a,b->m", G, u, u)
        assert np.array_equal(dq, u)
        assert np.allclose(du, ref, rtol=1e-7, atol=1e-9 * np.abs(ref).max()), (du, ref)


@pytest.mark.parametrize("a", [0.0, 0.9375, 0.998])
def test_icond_null_and_geometry(a):
    r_obs, i_obs, j_obs = 1000.0, 60.0, 30.0
    si, ci, sj, cj = O.camera_trig(i_obs, j_obs)
    for al, be in [(0.0, 0.0), (3.5, -2.0), (-31.5, 31.5)]:
        q, u = O.gr_icond(a, r_obs, i_obs, j_obs, al, be)
        k = O.down(a, q, u)
        assert abs(O.getuu(a, q, u)) < 1e-13 * abs(k[0])
        assert u[0] == 1.0
        n = np.array([si * cj, si * sj, ci])
        # spatial direction is the line of sight, scaled
        assert np.allclose(np.cross(u[1:], n), 0, atol=1e-15)
        # image-plane offset is perpendicular to the line of sight
        off = q[1:] - r_obs * n
        assert off @ n == pytest.approx(0.0, abs=1e-10)
        assert np.linalg.norm(off) == pytest.approx(math.hypot(al, be), rel=1e-12, abs=1e-12)


def conserved(a, s):
    k = O.down(a, s[:4], s[4:8])
    return -k[0], s[1] * k[2] - s[2] * k[1], O.getuu(a, s[:4], s[4:8])


def test_rk4_conserves_energy_angular_momentum():
    a = 0.9375
    p = O.make_params(a)
    q, u = O.gr_icond(a, 30.0, 75.0, 0.0, 6.0, 2.0)
    s = np.concatenate([q, u])
    E0, L0, _ = conserved(a, s)
    for _ in range(3000):
        s = O.integrate(p, s, -0.01)
    E1, L1, uu = conserved(a, s)
    assert np.linalg.norm(s[1:4]) < 25  # moved in
    assert E1 == pytest.approx(E0, rel=1e-9)
    assert L1 == pytest.approx(L0, rel=1e-9, abs=1e-10)
    assert abs(uu) < 1e-9


def test_rk4_fourth_order():
    a = 0.9375
    p = O.make_params(a)
    q, u = O.gr_icond(a, 20.0, 60.0, 0.0, 5.0, 1.0)
    s0 = np.concatenate([q, u])

    def run(h, n):
        s = s0.copy()
        for _ in range(n):
            s = O.integrate(p, s, h)
        return s

    ref = run(-0.0025, 1600)
    e1 = np.abs(run(-0.08, 50) - ref).max()
    e2 = np.abs(run(-0.04, 100) - ref).max()
    assert 12.0 < e1 / e2 < 20.0


def test_step_quantisation_known_answers():
    """driver.cl:81-99 with dt=-32, n_sub=1024: n = 2147482624, the far-field
    cap gives dh = 2097151 units = exactly -0.03125, 1024 steps a launch."""
    n_sub, dt = 1024, -32.0
    n = (2 ** 31 - 1) // n_sub * n_sub
    assert n == 2147482624
    assert int(dt / n_sub / dt * n) == 2097151
    assert 2097151 * dt / n == -0.03125
    # the 0.01 cap of the gridded path
    assert int(-0.01 / dt * n) == 671088
    assert 671088 * dt / n == pytest.approx(-0.0099999952316261447, rel=1e-16)
    p = O.make_params(0.9375)
    q, u = O.gr_icond(0.9375, 1000.0, 60.0, 0.0, 10.0, 10.0)
    s = np.concatenate([q, u])
    done = O.C.c_int()
    steps = O.lib().og_evolve_ray(O.C.byref(p), O._dp(s), dt, n_sub, O.C.byref(done))
    assert steps == 1024 and done.value == 0
    # the ray moved 32 units of affine parameter inward along the line of sight
    assert np.linalg.norm(s[1:4]) < 1000.0 - 31.0
    # the cap itself (phys.cl:61-63): 0.01 eps^2 below |dt/n_sub|
    q = ks_point(0.9375, 2.0, 1.0, 0.0)
    eps = 2.0 - p.r_h
    assert O.getdt(p, q, u, -0.03125) == pytest.approx(-0.01 * eps * eps, rel=1e-15)
    # the horizon stop and the outward escape stop return zero
    assert O.getdt(p, ks_point(0.9375, 1.01 * p.r_h * 0.999, 1.0, 0.0), u, -0.03125) == 0.0
    far = np.array([0.0, 1100.0, 0.0, 0.0])
    assert O.getdt(p, far, np.array([1.0, -1.0, 0, 0]), -0.03125) == 0.0     # outward
    assert O.getdt(p, far, np.array([1.0, 1.0, 0, 0]), -0.03125) == -0.03125  # inward
    pref = O.make_params(0.9375, escape=O.ESC_REF)
    assert O.getdt(pref, far, np.array([1.0, -1.0, 0, 0]), -0.03125) == -0.03125


def test_schwarzschild_critical_impact_parameter():
    p = O.make_params(0.0)
    bc = 3 * math.sqrt(3)
    for b, captured in [(bc - 0.05, True), (bc + 0.05, False)]:
        data, info = O.icond_all(p, 2, 1, 4 * b, 0.0, 1000.0, 90.0, 0.0)  # alpha = +-b
        for _ in range(80):
            O.evolve_all(p, data, info, -32.0, 1024)
        r = math.sqrt(O.getrr(0.0, data[1, :4]))
        assert (r < 1.02 * p.r_h) == captured and (captured or r > 999.0), (b, r)


def test_uniform_slab_transfer_closed_form():
    """n = 1, shift = 1: dtau/dl = -alpha, dI/dl = -j exp(-tau); over one
    launch of affine length 32: tau = 32 alpha, I = (j/alpha)(1 - e^-tau)."""
    nus = (1.0, 2.67, 4.0, 8.0)
    p = O.make_params(0.0, nu=nus, fluid=O.FLUID_UNIFORM, escape=O.ESC_REF)
    data, info = O.icond_all(p, 1, 1, 1.0, 1.0, 1000.0, 60.0, 0.0)
    O.evolve_all(p, data, info, -32.0, 1024)
    for i, nu in enumerate(nus):
        j, al = nu ** -0.5, 0.1 * nu ** -2.5
        tau = data[0, 8 + 4 + i]
        I = data[0, 8 + i]
        assert tau == pytest.approx(32 * al, rel=1e-13)
        assert I == pytest.approx(j / al * (1 - math.exp(-32 * al)), rel=1e-11)


def test_disk_transfer_channel_scaling():
    """tau_nu is a common path integral times nu^-2.5, exactly."""
    nus = (1.0, 2.67, 4.0, 8.0)
    a = 0.9375
    p = O.make_params(a, nu=nus)
    data, info = O.icond_all(p, 4, 4, 20.0, 20.0, 1000.0, 60.0, 0.0)
    for _ in range(40):
        O.evolve_all(p, data, info, -32.0, 1024)
    tau = data[:, 12:16] * np.array(nus) ** 2.5
    I = data[:, 8:12]
    assert (I > 0).any() and (tau[:, 0] > 0).any()
    assert np.allclose(tau, tau[:, :1], rtol=1e-12, atol=0)


def test_disk_flow_properties():
    a = 0.9375
    p = O.make_params(a, nu=(1.0,))
    # inside the ISCO: no emission
    n, sh = O.flow_disk(p, ks_point(a, 1.9, 1.5707963, 0.0), np.array([-1.0, 0, 0, 0]))
    assert n == 0.0
    # on the equator at r: n = r^-1.5, Keplerian shift for k_mu = (-1,0,0,0)
    r = 6.0
    q = ks_point(a, r, math.pi / 2, 0.3)
    n, sh = O.flow_disk(p, q, np.array([-1.0, 0, 0, 0]))
    assert n == pytest.approx(r ** -1.5, rel=1e-12)
    # u^t of a prograde circular equatorial orbit (Bardeen et al. 1972)
    ut = (r ** 1.5 + a) / math.sqrt(r ** 0.75 * r ** 0.75 * (r ** 1.5 - 3 * r ** 0.5 + 2 * a))
    assert sh == pytest.approx(ut, rel=1e-12)


def test_fused_and_unfused_oracles_agree():
    """The oracle fuses the geodesic RHS's and RK4's multiply-adds (as an
    OpenCL compiler may under FP_CONTRACT ON, sim-org/KS.cl:263-378,
    RK4.cl:37-47).  The same sources with every FMA split into a multiply
    and an add take the same steps and end within 1e-10 of each ray's state
    scale on a 64 M camera at r = 1000 (a = 0.9375, i = 60), so the
    choice is far inside the parity tolerance; the RHS itself differs by a
    few ulp only."""
    a = 0.9375
    d0, _, s0 = O.render(O.make_params(a), 24, 24, 64.0, 64.0, 1000.0, 60.0, 0.0)
    with O.unfused():
        d1, _, s1 = O.render(O.make_params(a), 24, 24, 64.0, 64.0, 1000.0, 60.0, 0.0)
        q = np.array([0.0, 3.1, -2.2, 1.7])
        u = np.array([1.0, -0.3, 0.5, 0.2])
        du1 = O.gr_rhs(a, q, u)[1]
    du0 = O.gr_rhs(a, q, u)[1]
    assert np.array_equal(s0, s1)
    assert not np.array_equal(d0[:, :8], d1[:, :8])      # the two builds differ
    scale = np.abs(d1[:, :8]).max(axis=1)
    assert (np.abs(d0[:, :8] - d1[:, :8]).max(axis=1) / scale).max() <= 1e-10
    assert np.abs(du0 - du1).max() <= 8 * np.finfo(float).eps * np.abs(du1).max()
