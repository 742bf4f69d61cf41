"""The reference's transfer physics (sim-org/rt.cl:28-198) and flow
(sim-org/flow.cl:36-58) in the oracle, pinned against the published
formulas they implement, evaluated here independently:

  ln K_2          scipy.special.kv (the table of rt.cl:51-68 is ln K_2(1/theta)
                  at theta = 0.1 * 1000^(i/60));
  B_nu            Planck's law via expm1, Rayleigh-Jeans and Wien limits;
  Gaunt factor    Rybicki & Lightman (1979) fig. 5.2 region formulas with
                  zeta = exp(Euler gamma);
  j_ff            Rybicki & Lightman eq. 5.14b / 4 pi, Z = 1, n_i = n_e;
  j_syn           Leung, Gammie & Noble (2011) eq. 72 with scipy's K_2;
  transfer        closed forms on a flat grid: with the fluid at rest and
                  u^t = 1, shift = 1, j and alpha = j / B are constant or
                  follow the density along straight rays, so I = int j dl and
                  tau = int alpha dl (RK4 is exact for the quadratic n_e^2 of a
                  linear density).
"""
import math

import numpy as np
import pytest
from scipy import special

from oracle import oracle as O

# cgs constants (CODATA 2010)
C_LIGHT, H, G_N, K_B = 2.99792458e10, 6.62606957e-27, 6.67384800e-8, 1.38064881e-16
RY, Q_E, M_E, M_SUN = 2.17987197e-11, 4.80320425e-10, 9.10938291e-28, 1.98910000e33
L_UNIT = G_N * M_SUN / C_LIGHT ** 2       # cm per M_sun of G M / c^2
M_ADM = 4.3e6

RT = O.Rt()


@pytest.mark.parametrize("x", np.logspace(-3, 2, 23))
def test_log_k2_matches_scipy(x):
    assert O.log_k2(x) == pytest.approx(math.log(special.kv(2, x)), rel=1e-13, abs=1e-13)


def test_log_k2_table_nodes_and_interpolation():
    for i in (0, 1, 17, 59, 60):
        theta = 0.1 * 1000.0 ** (i / 60)
        want = math.log(special.kv(2, 1 / theta))
        assert RT.log_k2it(theta) == pytest.approx(want, rel=1e-12, abs=1e-12)
    # linear in ln(theta) between nodes
    t0, t1 = 0.1 * 1000.0 ** (30 / 60), 0.1 * 1000.0 ** (31 / 60)
    tm = math.sqrt(t0 * t1)
    assert RT.log_k2it(tm) == pytest.approx(0.5 * (RT.log_k2it(t0) + RT.log_k2it(t1)), rel=1e-12)
    # linear interpolation in ln(theta) over cells of width ln(1000)/60: the
    # error is about width^2/8 times the curvature, ~1/theta for small theta
    for theta in np.logspace(-1, 2, 37)[1:-1] * 1.013:
        assert abs(RT.log_k2it(theta) - math.log(special.kv(2, 1 / theta))) < 2e-3 / theta


@pytest.mark.parametrize("nu,te", [(1e9, 1.0), (1e11, 1e12), (1e15, 0.1), (3e20, 10.0),
                                   (2.3e11, 1e-3), (5e14, 3e-3)])
def test_planck(nu, te):
    x = H * nu / (M_E * C_LIGHT ** 2 * te)
    want = 2 * H * nu ** 3 / C_LIGHT ** 2 / np.expm1(x)
    assert RT.b_planck(nu, te) == pytest.approx(want, rel=1e-12)
    if x < 1e-6:   # Rayleigh-Jeans: 2 nu^2 k T / c^2
        assert RT.b_planck(nu, te) == pytest.approx(2 * nu ** 2 * M_E * te, rel=1e-5)
    if x > 30:     # Wien
        assert RT.b_planck(nu, te) == pytest.approx(2 * H * nu ** 3 / C_LIGHT ** 2 * math.exp(-x),
                                                    rel=1e-12)


def gaunt_rl(x, y):
    """Region formulas: x = kT/Ry, y = h nu/kT; zeta = e^gamma."""
    zeta = math.exp(np.euler_gamma)
    if x > 1:
        if y > 1:
            return math.sqrt(3 / (math.pi * y))
        return math.sqrt(3) / math.pi * math.log(4 / (zeta * y))
    if x * y > 1:
        return math.sqrt(12 / (x * y))
    if y > math.sqrt(x):
        return 1.0
    return max(math.sqrt(3) / math.pi * math.log(4 / zeta ** 2.5 * math.sqrt(x) / y), 1e-28)


@pytest.mark.parametrize("x,y", [(4e16, 8e-15), (10.0, 3.0), (5.0, 0.2), (0.5, 4.0),
                                 (0.25, 0.7), (0.04, 0.01), (0.81, 0.5), (1e-4, 1e-6)])
def test_gaunt_regions(x, y):
    assert RT.gaunt(x, y) == pytest.approx(gaunt_rl(x, y), rel=1e-11)


def test_gaunt_guard_at_zero_frequency():
    # rt.cl adds eps = 1e-28 to y inside the logarithm: finite at y = 0
    zeta = math.exp(np.euler_gamma)
    assert RT.gaunt(4e16, 0.0) == pytest.approx(math.sqrt(3) / math.pi * math.log(4 / (zeta * 1e-28)),
                                                rel=1e-12)


@pytest.mark.parametrize("nu,te,ne", [(1e11, 1e12, 0.5), (2.67e11, 1e12, 1.0), (1e12, 3.0, 2.0),
                                      (5e18, 10.0, 1e-3)])
def test_bremsstrahlung_rybicki_lightman(nu, te, ne):
    T = te * M_E * C_LIGHT ** 2 / K_B
    g = gaunt_rl(K_B * T / RY, H * nu / (K_B * T))
    eps = 6.8e-38 * ne * ne / math.sqrt(T) * math.exp(-H * nu / (K_B * T)) * g
    want = M_ADM * L_UNIT * eps / (4 * math.pi)
    assert RT.j_ff(nu, te, ne) == pytest.approx(want, rel=1e-11)


@pytest.mark.parametrize("te", [0.5, 3.0, 40.0, 250.0, 1e4])
def test_synchrotron_leung_gammie_noble(te):
    nu, ne, b, cos_t = 2.3e11, 1e5, 30.0, 0.4
    nu_c = Q_E * b / (2 * math.pi * M_E * C_LIGHT)
    nu_s = 2 / 9 * nu_c * te * te * math.sqrt(1 - cos_t ** 2)
    X = nu / nu_s
    j = (ne * Q_E ** 2 * nu_s * math.sqrt(2) * math.pi / (3 * special.kv(2, 1 / te) * C_LIGHT)
         * (X ** 0.5 + 2 ** (11 / 12) * X ** (1 / 6)) ** 2 * math.exp(-X ** (1 / 3)))
    want = M_ADM * L_UNIT * j
    tol = 2e-3 / te if te <= 100 else 1e-6   # table interpolation; asymptotic K_2 above it
    assert RT.j_syn(nu, te, ne, b, cos_t) == pytest.approx(want, rel=tol)


def test_synchrotron_vanishes_for_the_reference_flow():
    # flow.cl:52-55: b = 0 and bkcos = 0 give exactly zero synchrotron emission
    for nu in (1e9, 2.67e11, 8e11):
        assert RT.j_syn(nu, 1e12, 0.7, 0.0, 0.0) == 0.0
    assert RT.j_syn(1e11, 0.05, 1.0, 10.0, 0.3) == 0.0      # below T_MIN
    assert RT.j_syn(1e11, 5.0, 1.0, 10.0, 1.0) == 0.0       # along the field


def test_transfer_rhs_structure():
    nu = [1e11, 2.67e11, 8e11]
    I, tau = np.array([1e-30, 0.0, 2e-31]), np.array([0.0, 0.3, 1.5])
    ne, te, shift = 0.8, 1e12, 1.3
    dI, dtau = RT.thermal(nu, ne, te, 0.0, 0.0, shift, I, tau)
    for i, n in enumerate(nu):
        L = RT.j_ff(n * shift, te, ne)
        assert dI[i] == pytest.approx(-L * math.exp(-tau[i]) / shift ** 2, rel=1e-14)
        assert dtau[i] == pytest.approx(-L * shift / RT.b_planck(n * shift, te), rel=1e-14)


def flat_grid(n=9, lo=-10.0, hi=10.0, rho_fn=None):
    xs = np.linspace(lo, hi, n)
    X, Y, Z = np.meshgrid(xs, xs, xs, indexing="ij")
    X, Y, Z = (A.transpose(2, 1, 0).ravel() for A in (X, Y, Z))  # x fastest
    gam = np.zeros((X.size, 40))
    met = np.zeros((X.size, 10))
    met[:, 0], met[:, 4], met[:, 7], met[:, 9] = -1.0, 1.0, 1.0, 1.0
    rho = rho_fn(X, Y, Z) if rho_fn else np.full(X.size, 0.6)
    return (n,) * 3, (lo,) * 3, (hi,) * 3, gam, met, rho


def run_rays(G, rays, dt=-4.0, n_sub=64, n_dump=2):
    data = np.array(rays, dtype=np.float64)
    steps = np.zeros(len(data), dtype=np.int64)
    for _ in range(n_dump):
        G.evolve_all(data, np.zeros(len(data)), dt, n_sub, steps, 1)
    return data


def test_flat_slab_closed_form():
    nu = (1e11, 2.67e11, 4e11, 8e11)
    n, lo, hi, gam, met, rho = flat_grid()
    G = O.GridSpacetime(n, lo, hi, [gam], [met], rho=[rho], nu=nu)
    ne, shift = G.flow([0.0, 1.0, 2.0, 3.0], [1.0, 0.6, 0.0, -0.8])
    assert ne == pytest.approx(0.6, rel=1e-15) and shift == pytest.approx(1.0, rel=1e-15)
    s0 = np.zeros(G.n_data)
    s0[:8] = [0.0, 1.0, 2.0, 3.0, 1.0, 0.6, 0.0, -0.8]
    (s,) = run_rays(G, [s0])
    L = -s[0]                       # u^t = 1: affine length = coordinate time
    assert L == pytest.approx(8.0, rel=1e-14)
    np.testing.assert_allclose(s[1:4], s0[1:4] - L * s0[5:8], rtol=1e-13)
    for i, v in enumerate(nu):
        j = RT.j_ff(v, 1e12, 0.6)
        B = RT.b_planck(v, 1e12)
        assert s[8 + i] == pytest.approx(j * L, rel=1e-12)          # e^-tau = 1 to 1e-30
        assert s[12 + i] == pytest.approx(j / B * L, rel=1e-12)


def test_flat_slab_with_field_closed_form():
    """A uniform field b^mu = (0, 0, 0, B) (flow.cl:47 with B instead of 0):
    in flat space with the fluid at rest, bkcos = k_z B / (shift B + eps)
    (flow.cl:55) is the cosine between the ray and z, constant along a
    straight ray, so I = (j_ff + j_syn(cos)) L and tau = I / B_nu."""
    nu = (1e11, 8e11)
    B = 20.0
    n, lo, hi, gam, met, rho = flat_grid()
    G = O.GridSpacetime(n, lo, hi, [gam], [met], rho=[rho], nu=nu, b_field=B)
    u = [1.0, 0.6, 0.0, -0.8]
    ne, shift, bkcos = G.flow([0.0, 1.0, 2.0, 3.0], u, with_bkcos=True)
    assert shift == pytest.approx(1.0, rel=1e-15) and bkcos == pytest.approx(-0.8, rel=1e-15)
    assert G.flow([0.0, 1.0, 2.0, 3.0], [1.0, 0.0, 0.0, 1.0], with_bkcos=True)[2] == \
        pytest.approx(1.0, rel=1e-15)
    s0 = np.zeros(G.n_data)
    s0[:8] = [0.0, 1.0, 2.0, 3.0] + u
    (s,) = run_rays(G, [s0])
    L = 8.0
    for i, v in enumerate(nu):
        js = RT.j_syn(v, 1e12, 0.6, B, -0.8)
        j = RT.j_ff(v, 1e12, 0.6) + js
        assert js > 1e10 * RT.j_ff(v, 1e12, 0.6)        # the field's emission dominates
        assert s[8 + i] == pytest.approx(j * L, rel=1e-12)
        assert s[10 + i] == pytest.approx(j / RT.b_planck(v, 1e12) * L, rel=1e-12)
    # B = 0 is the reference flow: bremsstrahlung alone, bkcos = 0
    G0 = O.GridSpacetime(n, lo, hi, [gam], [met], rho=[rho], nu=nu)
    assert G0.flow([0.0, 1.0, 2.0, 3.0], u, with_bkcos=True)[2] == 0.0
    (s,) = run_rays(G0, [s0])
    assert s[8] == pytest.approx(RT.j_ff(nu[0], 1e12, 0.6) * L, rel=1e-12)


def test_flat_density_gradient_quadratic_emission():
    nu = (2.67e11,)
    n, lo, hi, gam, met, rho = flat_grid(rho_fn=lambda X, Y, Z: 1.0 + 0.05 * X - 0.02 * Z)
    G = O.GridSpacetime(n, lo, hi, [gam], [met], rho=[rho], nu=nu)
    s0 = np.zeros(G.n_data)
    s0[:8] = [0.0, -2.0, 0.5, 1.0, 1.0, 0.8, 0.0, 0.6]
    (s,) = run_rays(G, [s0])
    j1 = RT.j_ff(nu[0], 1e12, 1.0)   # j = j1 n_e^2
    L = 8.0

    def ne(l):  # density at affine distance l back along the ray
        x, z = s0[1] - l * s0[5], s0[3] - l * s0[7]
        return 1.0 + 0.05 * x - 0.02 * z
    # Simpson is exact for the quadratic n_e^2 along the ray
    want = j1 * L / 6 * (ne(0) ** 2 + 4 * ne(L / 2) ** 2 + ne(L) ** 2)
    assert s[8] == pytest.approx(want, rel=1e-12)


def test_grid_without_density_rejects_transfer():
    n, lo, hi, gam, met, rho = flat_grid()
    with pytest.raises(ValueError):
        O.GridSpacetime(n, lo, hi, [gam], [met], nu=(1e11,))
