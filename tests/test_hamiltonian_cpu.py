"""The FP32 form's geodesic equations (gray_device.h ks_ham) against the
oracle's Kerr-Schild geodesic RHS (sim-org/KS.cl:263-378).

The FP32 form carries the covariant momentum p_mu = g_{mu nu} u^nu and
integrates Hamilton's equations for H = g^{mu nu} p_mu p_nu / 2 with
g^{mu nu} = eta^{mu nu} - f l^mu l^nu, l^mu = (-1, lx, ly, lz):

    u^mu = eta^{mu nu} p_nu - f L l^mu,                  L = l^mu p_mu
    dp_i/dlambda = L (L d_i f / 2 + f (d_i l^nu) p_nu),  dp_t/dlambda = 0.

They describe the same curves as the oracle's u^mu form exactly when, along
the oracle's flow, dp_mu/dlambda = d(g_{mu nu} u^nu)/dlambda =
u^k d_k(g_{mu nu}) u^nu + g_{mu nu} du^nu/dlambda.  This test evaluates both
sides at random events outside the horizon, with d_k(g u) and d_i f, d_i l
by five-point central differences of closed forms that do not come from the device
code (f = 2 r^3 / (r^4 + a^2 z^2), l from the Kerr-Schild r), and checks
the velocity map against the oracle's down().  A sign, a factor 1/2 or a
dropped term in the Hamiltonian RHS fails it."""
import numpy as np
import pytest

from oracle import oracle as O


def ks_r(a, x, y, z):
    """Kerr-Schild r: the positive root of r^4 - (R^2 - a^2) r^2 - a^2 z^2."""
    k = 0.5 * (x * x + y * y + z * z - a * a)
    return np.sqrt(np.sqrt(k * k + a * a * z * z) + k)


def f_l(a, x):
    r = ks_r(a, x[0], x[1], x[2])
    f = 2.0 * r ** 3 / (r ** 4 + a * a * x[2] ** 2)
    l = np.array([(r * x[0] + a * x[1]) / (r * r + a * a),
                  (r * x[1] - a * x[0]) / (r * r + a * a), x[2] / r])
    return f, l


H = 1e-3


def deriv(fn, x, e):
    """Five-point central difference of fn along e (|e| = H)."""
    return (-fn(x + 2 * e) + 8 * fn(x + e) - 8 * fn(x - e) + fn(x - 2 * e)) / (12 * H)


def ham_rhs(a, q, p, sign=1.0):
    """Hamilton's equations above (sign = -1 flips the f (d_i l) p term)."""
    x = np.asarray(q[1:], dtype=float)
    f, l = f_l(a, x)
    L = -p[0] + l @ p[1:]
    u = np.concatenate(([-p[0] + f * L], p[1:] - f * L * l))
    dp = np.zeros(4)
    for i in range(3):
        e = np.eye(3)[i] * H
        df = deriv(lambda y: f_l(a, y)[0], x, e)
        dl = deriv(lambda y: f_l(a, y)[1], x, e)
        dp[1 + i] = L * (0.5 * L * df + sign * f * (dl @ p[1:]))
    return u, dp


def chain_rhs(a, q, u):
    """d(g u)/dlambda along the oracle's flow: u^k d_k(g) u + g du/dlambda."""
    _, du = O.gr_rhs(a, q, u)
    dgu = np.zeros(4)
    for k in range(1, 4):
        dgu += u[k] * deriv(lambda y: O.down(a, y, u), np.asarray(q, dtype=float),
                            np.eye(4)[k] * H)
    return dgu + O.down(a, q, du), du


def events(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        a = rng.uniform(-0.99, 0.99)
        q = np.concatenate(([rng.uniform(-50, 50)], rng.uniform(-30, 30, 3)))
        if ks_r(a, *q[1:]) < 2.5:
            continue
        u = np.concatenate(([1.0], rng.normal(size=3)))
        out.append((a, q, u))
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_velocity_map_inverts_down(seed):
    for a, q, u in events(40, seed):
        p = O.down(a, q, u)
        u_h, _ = ham_rhs(a, q, p)
        np.testing.assert_allclose(u_h, u, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_momentum_rhs_matches_the_geodesic_rhs(seed):
    for a, q, u in events(40, seed):
        p = O.down(a, q, u)
        dp_chain, du = chain_rhs(a, q, u)
        _, dp_h = ham_rhs(a, q, p)
        scale = np.abs(dp_chain).max() + np.abs(du).max()
        assert abs(dp_chain[0]) <= 1e-9 * scale          # p_t is conserved
        np.testing.assert_allclose(dp_h, dp_chain, rtol=0, atol=1e-9 * scale)


def test_a_planted_sign_error_fails():
    """The comparison above is sharp enough to see a flipped term."""
    worst = 0.0
    for a, q, u in events(10, 7):
        p = O.down(a, q, u)
        dp_chain, du = chain_rhs(a, q, u)
        _, bad = ham_rhs(a, q, p, sign=-1.0)
        scale = np.abs(dp_chain).max() + np.abs(du).max()
        worst = max(worst, np.abs(bad - dp_chain).max() / scale)
    assert worst > 1e-3


def closed_form(a, q, p):
    """The device's closed-form evaluation (gray_device.h ks_ham), in
    FP64: d_i(l.p) = d_i r K1 + H_i with K1 = (A - 2 r S_xy) / (r^2 + a^2) -
    l_z p_z / r, d_i f = F1 d_i r (+ F2 for z), folded into dp_i."""
    x, y, z = q[1:]
    pt, px, py, pz = p
    aa, zz = a * a, z * z
    kk = 0.5 * ((x * x + y * y + zz) - aa)
    dd = np.sqrt(kk * kk + aa * zz)
    rr = dd + kk
    r = np.sqrt(rr)
    ir = 1 / r
    half_rdd = 0.5 / (r * dd)
    two_rq = 2 * rr / (rr * rr + aa * zz)
    ss = rr + aa
    iss = 1 / ss
    f = two_rq * r
    F1 = two_rq * (3 - 2 * rr * two_rq)
    F2 = -two_rq * (two_rq * aa * z) * ir
    lx, ly, lz = iss * (x * r + a * y), iss * (y * r - a * x), z * ir
    A, B = x * px + y * py, y * px - x * py
    s_xy = iss * (r * A + a * B)
    L = s_xy + lz * pz - pt
    fL = f * L
    u = np.array([fL - pt, px - fL * lx, py - fL * ly, pz - fL * lz])
    K1 = iss * (A - 2 * r * s_xy) - lz * pz * ir
    M = 0.5 * L * F1 + f * K1
    hx, hy, hz = iss * (r * px - a * py), iss * (r * py + a * px), pz * ir
    cxy, cz = half_rdd * rr, half_rdd * ss
    dp = np.array([0.0, L * (x * cxy * M + f * hx), L * (y * cxy * M + f * hy),
                   L * (z * cz * M + f * hz + 0.5 * L * F2)])
    return u, dp


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_closed_form_matches_hamiltons_equations(seed):
    for a, q, u in events(40, seed):
        p = O.down(a, q, u)
        u_h, dp_h = ham_rhs(a, q, p)
        u_c, dp_c = closed_form(a, q, p)
        np.testing.assert_allclose(u_c, u_h, rtol=1e-13, atol=1e-13)
        scale = np.abs(dp_h).max()
        np.testing.assert_allclose(dp_c, dp_h, rtol=0, atol=1e-9 * scale)
