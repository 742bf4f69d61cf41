"""Boyer-Lindquist view of Kerr-Schild states in the oracle (restating
sim-org/KS.cl:63-149), pinned by:
  * the spheroidal map: a point built from (r, theta, phi) with
    x + i y = (r + i a) e^{i phi} sin(theta), z = r cos(theta) maps back to
    (r, theta, phi); in this convention the metric of KS.cl:100-108 takes the
    Boyer-Lindquist form after r-only shifts of t and phi (checked below);
  * invariance: g_BL(u_BL, u_BL) with the closed-form Boyer-Lindquist metric
    equals g_KS(u, u) for arbitrary (also non-null) vectors, which fixes the
    three Jacobian rows and the t and phi corrections;
  * consistency: along an RK4 geodesic, the finite-difference derivative of
    the BL position map equals the BL velocity (this is where KS.cl:146,149,
    which use u^phi and u^t instead of u^r, would fail);
  * a = 0: t_BL = t_KS - 2 ln((r - 2) / (r_match - 2)), phi unchanged.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def ks_point(a, r, th, ph, t=0.3):
    return np.array([t, (r * math.cos(ph) - a * math.sin(ph)) * math.sin(th),
                     (r * math.sin(ph) + a * math.cos(ph)) * math.sin(th), r * math.cos(th)])


def g_bl(a, r, th):
    s2, c2 = math.sin(th) ** 2, math.cos(th) ** 2
    sig, dl = r * r + a * a * c2, r * r - 2 * r + a * a
    g = np.zeros((4, 4))
    g[0, 0] = -(1 - 2 * r / sig)
    g[0, 3] = g[3, 0] = -2 * a * r * s2 / sig
    g[3, 3] = (r * r + a * a + 2 * a * a * r * s2 / sig) * s2
    g[1, 1] = sig / dl
    g[2, 2] = sig
    return g


@pytest.mark.parametrize("a", [0.0, 0.5, 0.9375, 0.998])
def test_spheroidal_map(a):
    for r, th, ph in [(3.0, 0.7, 0.4), (10.0, 2.1, -2.5), (1000.0, 1.2, 3.0), (2.5, 0.3, 1.6)]:
        s = O.getsphks(a, ks_point(a, r, th, ph))
        assert s[0] == 0.3
        np.testing.assert_allclose(s[1:], [r, th, ph], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("a", [0.0, 0.6, 0.9375])
def test_metric_invariance(a):
    rng = np.random.default_rng(8)
    for _ in range(30):
        r = rng.uniform(2.2, 30.0)
        q = ks_point(a, r, rng.uniform(0.2, 2.9), rng.uniform(-3, 3))
        u = rng.normal(size=4)
        ub = O.getblu(a, q, u)
        rr, th = O.getsphks(a, q)[1:3]
        lhs = ub @ g_bl(a, rr, th) @ ub
        rhs = O.getuu(a, q, u)
        assert lhs == pytest.approx(rhs, rel=1e-9, abs=1e-9 * np.abs(u).max() ** 2)


def test_bl_velocity_is_the_derivative_of_the_bl_position():
    a, rm = 0.9375, 50.0
    p = O.make_params(a)
    q, u = O.gr_icond(a, 40.0, 1.0, 0.0, 3.0, -2.0)
    s = np.concatenate([q, u])
    for _ in range(300):              # move in to r ~ 25
        s = O.integrate(p, s, -0.05)
    assert 20.0 < O.getsphks(a, s[:4])[1] < 35.0
    h = 1e-4
    sp, sm = O.integrate(p, s, h), O.integrate(p, s, -h)
    dbl = (O.getbl(a, rm, sp[:4]) - O.getbl(a, rm, sm[:4])) / (2 * h)
    ub = O.getblu(a, s[:4], s[4:])
    np.testing.assert_allclose(dbl, ub, rtol=1e-7, atol=1e-9)


def test_schwarzschild_time_shift():
    q = ks_point(0.0, 7.0, 1.0, 0.5, t=4.0)
    b = O.getbl(0.0, 20.0, q)
    assert b[0] == pytest.approx(4.0 - 2 * math.log(5.0 / 18.0), rel=1e-14)
    assert b[3] == pytest.approx(0.5, rel=1e-14)
