"""ctypes binding of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point as the
checker and by the benchmark's CPU-baseline leg.  The product path never
imports this module.
"""
import contextlib
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# The oracle as built (geodesic RHS and RK4 stages with fused multiply-adds)
# and the same sources built with every FMA as a multiply and an add.
_LIBS = {"fused": os.path.join(_HERE, "liboracle.so"),
         "unfused": os.path.join(_HERE, "liboracle_unfused.so")}

MAX_FREQ = 8
ESC_REF, ESC_OUTWARD = 0, 1
FLUID_NONE, FLUID_DISK, FLUID_UNIFORM = 0, 1, 2


class Params(C.Structure):
    _fields_ = [
        ("a", C.c_double),
        ("r_h", C.c_double),
        ("eps_stop", C.c_double),
        ("escape", C.c_int),
        ("r_esc", C.c_double),
        ("n_freq", C.c_int),
        ("nu", C.c_double * MAX_FREQ),
        ("fluid", C.c_int),
        ("r_isco", C.c_double),
    ]


_libs = {}
_variant = "fused"


@contextlib.contextmanager
def unfused():
    """Within this block every oracle call runs the unfused build."""
    global _variant
    old, _variant = _variant, "unfused"
    try:
        yield
    finally:
        _variant = old


def lib():
    L = _libs.get(_variant)
    if L is None:
        path = _LIBS[_variant]
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        pp = C.POINTER(Params)
        L.og_setup.argtypes = [pp]
        L.og_r_isco.argtypes = [C.c_double]
        L.og_r_isco.restype = C.c_double
        L.og_getrr.argtypes = [C.c_double, dp]
        L.og_getrr.restype = C.c_double
        L.og_geteps.argtypes = [C.c_double, dp]
        L.og_geteps.restype = C.c_double
        L.og_down.argtypes = [C.c_double, dp, dp, dp]
        L.og_getuu.argtypes = [C.c_double, dp, dp]
        L.og_getuu.restype = C.c_double
        L.og_gr_icond.argtypes = [C.c_double] * 8 + [dp, dp]
        L.og_gr_rhs.argtypes = [C.c_double, dp, dp, dp, dp]
        L.og_getdt.argtypes = [pp, dp, dp, C.c_double]
        L.og_getdt.restype = C.c_double
        L.og_flow_disk.argtypes = [pp, dp, dp, dp, dp]
        L.og_rhs.argtypes = [pp, dp, dp]
        L.og_integrate.argtypes = [pp, dp, C.c_double]
        L.og_icond_ray.argtypes = [pp, C.c_size_t, C.c_size_t] + [C.c_double] * 7 + [C.c_size_t, C.c_size_t, dp]
        L.og_evolve_ray.argtypes = [pp, dp, C.c_double, C.c_size_t, C.POINTER(C.c_int)]
        L.og_evolve_ray.restype = C.c_int64
        L.og_icond_all.argtypes = [pp, C.c_size_t, C.c_size_t] + [C.c_double] * 7 + [dp, dp]
        L.og_evolve_all.argtypes = [pp, C.c_size_t, dp, dp, C.c_double, C.c_size_t,
                                    C.POINTER(C.c_int64), C.c_int]
        L.og_evolve_all.restype = C.c_int64
        L.og_evolve_subset.argtypes = [pp, C.c_size_t, C.c_size_t, dp, C.c_double,
                                       C.c_size_t, C.c_size_t, C.c_int]
        L.og_evolve_subset.restype = C.c_int64
        _libs[_variant] = L
    return L


def _dp(x):
    return x.ctypes.data_as(C.POINTER(C.c_double))


def _vec(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def make_params(a, escape=ESC_OUTWARD, eps_stop=None, r_esc=1000.0, nu=(), fluid=None):
    """Parameters of one render.  eps_stop defaults to 0.01*r_h for the
    'r < 1.01 r_H' horizon rule (ESC_OUTWARD) and 0.01 for the reference
    rule (ESC_REF, sim-org/phys.cl:54)."""
    p = Params()
    p.a = a
    lib().og_setup(C.byref(p))
    if eps_stop is None:
        eps_stop = 0.01 * p.r_h if escape == ESC_OUTWARD else 0.01
    p.eps_stop = eps_stop
    p.escape = escape
    p.r_esc = r_esc
    p.n_freq = len(nu)
    for i, v in enumerate(nu):
        p.nu[i] = v
    if fluid is None:
        fluid = FLUID_DISK if len(nu) else FLUID_NONE
    p.fluid = fluid
    p.r_isco = lib().og_r_isco(a)
    return p


def camera_trig(i_obs, j_obs):
    """sin/cos of the camera angles in degrees (sim-org/KS.cl:195-197),
    computed once per image on the host."""
    d2r = math.pi / 180.0
    return (math.sin(d2r * i_obs), math.cos(d2r * i_obs),
            math.sin(d2r * j_obs), math.cos(d2r * j_obs))


def getrr(a, q):
    return lib().og_getrr(a, _dp(_vec(q)))


def geteps(a, q):
    return lib().og_geteps(a, _dp(_vec(q)))


def down(a, q, u):
    k = np.zeros(4)
    lib().og_down(a, _dp(_vec(q)), _dp(_vec(u)), _dp(k))
    return k


def getuu(a, q, u):
    return lib().og_getuu(a, _dp(_vec(q)), _dp(_vec(u)))


def gr_icond(a, r_obs, i_obs, j_obs, alpha, beta):
    si, ci, sj, cj = camera_trig(i_obs, j_obs)
    q = np.zeros(4)
    u = np.zeros(4)
    lib().og_gr_icond(a, r_obs, si, ci, sj, cj, alpha, beta, _dp(q), _dp(u))
    return q, u


def gr_rhs(a, q, u):
    dq = np.zeros(4)
    du = np.zeros(4)
    lib().og_gr_rhs(a, _dp(_vec(q)), _dp(_vec(u)), _dp(dq), _dp(du))
    return dq, du


def getdt(p, q, u, dt):
    return lib().og_getdt(C.byref(p), _dp(_vec(q)), _dp(_vec(u)), dt)


def _bl_set():
    L = lib()
    dp = C.POINTER(C.c_double)
    L.og_getsphks.argtypes = [C.c_double, dp, dp]
    L.og_getbl.argtypes = [C.c_double, C.c_double, dp, dp]
    L.og_getsphksu.argtypes = [C.c_double, dp, dp, dp]
    L.og_getblu.argtypes = [C.c_double, dp, dp, dp]
    return L


def getsphks(a, q):
    out = np.zeros(4)
    _bl_set().og_getsphks(a, _dp(_vec(q)), _dp(out))
    return out


def getbl(a, r_match, q):
    out = np.zeros(4)
    _bl_set().og_getbl(a, r_match, _dp(_vec(q)), _dp(out))
    return out


def getsphksu(a, q, u):
    out = np.zeros(4)
    _bl_set().og_getsphksu(a, _dp(_vec(q)), _dp(_vec(u)), _dp(out))
    return out


def getblu(a, q, u):
    out = np.zeros(4)
    _bl_set().og_getblu(a, _dp(_vec(q)), _dp(_vec(u)), _dp(out))
    return out


def flow_disk(p, q, k):
    n = C.c_double()
    sh = C.c_double()
    lib().og_flow_disk(C.byref(p), _dp(_vec(q)), _dp(_vec(k)), C.byref(n), C.byref(sh))
    return n.value, sh.value


def rhs(p, s):
    s = _vec(s)
    ds = np.zeros_like(s)
    lib().og_rhs(C.byref(p), _dp(s), _dp(ds))
    return ds


def integrate(p, s, dt):
    s = _vec(s).copy()
    lib().og_integrate(C.byref(p), _dp(s), dt)
    return s


def n_data(p):
    return 8 + 2 * p.n_freq


def icond_all(p, w_rays, h_rays, w_img, h_img, r_obs, i_obs, j_obs):
    """AoS state [h_rays*w_rays, n_data] and info = g(u,u) per ray."""
    si, ci, sj, cj = camera_trig(i_obs, j_obs)
    n = w_rays * h_rays
    data = np.zeros((n, n_data(p)))
    info = np.zeros(n)
    lib().og_icond_all(C.byref(p), w_rays, h_rays, w_img, h_img, r_obs,
                       si, ci, sj, cj, _dp(data), _dp(info))
    return data, info


def icond_pixels(p, w_rays, h_rays, w_img, h_img, r_obs, i_obs, j_obs, pixels):
    """AoS state of the listed pixels (gi, gj) of a w_rays x h_rays image,
    og_icond_ray per pixel (for images too large to set up whole)."""
    si, ci, sj, cj = camera_trig(i_obs, j_obs)
    data = np.zeros((len(pixels), n_data(p)))
    for k, (gi, gj) in enumerate(pixels):
        row = data[k]
        lib().og_icond_ray(C.byref(p), w_rays, h_rays, w_img, h_img, r_obs, si, ci, sj, cj,
                           gi, gj, _dp(row))
    return data


def evolve_all(p, data, info, dt, n_sub, steps=None, n_threads=0):
    """One launch of sim-org/driver.cl:70-108 over every ray (in place)."""
    assert data.flags.c_contiguous and data.dtype == np.float64
    n = data.shape[0]
    sp = None
    if steps is not None:
        assert steps.dtype == np.int64 and steps.shape == (n,)
        sp = steps.ctypes.data_as(C.POINTER(C.c_int64))
    return lib().og_evolve_all(C.byref(p), n, _dp(data), _dp(info), dt, n_sub, sp, n_threads)


def render(p, w_rays, h_rays, w_img, h_img, r_obs, i_obs, j_obs,
           dt_dump=-32.0, n_sub=1024, n_dump=128, n_threads=0):
    """icond + n_dump launches; returns (data, info, steps per ray)."""
    data, info = icond_all(p, w_rays, h_rays, w_img, h_img, r_obs, i_obs, j_obs)
    steps = np.zeros(w_rays * h_rays, dtype=np.int64)
    for _ in range(n_dump):
        evolve_all(p, data, info, dt_dump, n_sub, steps, n_threads)
    return data, info, steps


def evolve_subset(p, data, stride, dt, n_sub, n_dump, n_threads=0):
    assert data.flags.c_contiguous and data.dtype == np.float64
    return lib().og_evolve_subset(C.byref(p), data.shape[0], stride, _dp(data), dt,
                                  n_sub, n_dump, n_threads)


# ---------------------------------------------------------------------------
# numerical-spacetime oracle (grid_oracle.c)

MAP_UNIFORM, MAP_FISHEYE = 0, 1


RT_T_GRID = 60
RT_TE_FLOW = 1e12
M_ADM_DEFAULT = 4.3e6


class RtConsts(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "planck1", "planck2", "c", "gaunt_a", "gaunt_b", "gaunt_c", "gaunt_d", "gaunt_e",
        "ff_x", "ff_y", "ff_f", "syn_nus", "syn_f", "syn_k", "k2_h")] + [
        ("log_k2", C.c_double * (RT_T_GRID + 1))]


class Grid(C.Structure):
    _fields_ = [
        ("n", C.c_int * 3),
        ("lo", C.c_double * 3),
        ("hi", C.c_double * 3),
        ("map", C.c_int),
        ("t1", C.c_double),
        ("t2", C.c_double),
        ("t_off", C.c_double),
        ("raw_icond", C.c_int),
        ("ref_rule", C.c_int),
        ("gam", C.c_void_p * 2),
        ("met", C.c_void_p * 2),
        ("rho", C.c_void_p * 2),
        ("n_freq", C.c_int),
        ("nu", C.c_double * MAX_FREQ),
        ("m_adm", C.c_double),
        ("rt", C.POINTER(RtConsts)),
        ("b_field", C.c_double),
    ]


_rt_set = set()   # variants whose argtypes are set


def _rlib():
    L = lib()
    if _variant not in _rt_set:
        rp, dp, d = C.POINTER(RtConsts), C.POINTER(C.c_double), C.c_double
        L.og_log_k2.argtypes = [d]
        L.og_log_k2.restype = d
        L.og_rt_consts_init.argtypes = [rp]
        for name, n in (("og_log_k2it", 1), ("og_b_planck", 2), ("og_gaunt", 2), ("og_j_ff", 4),
                        ("og_j_syn", 6)):
            fn = getattr(L, name)
            fn.argtypes = [rp] + [d] * n
            fn.restype = d
        L.og_rt_thermal.argtypes = [rp, C.c_int, dp, d, d, d, d, d, d, dp, dp, dp, dp]
        _rt_set.add(_variant)
    return L


def log_k2(x):
    """ln K_2(x) by quadrature (the oracle's table generator)."""
    return _rlib().og_log_k2(x)


class Rt:
    """The reference's transfer physics (rt.cl) with its constants."""

    def __init__(self):
        self.c = RtConsts()
        _rlib().og_rt_consts_init(C.byref(self.c))

    def log_k2it(self, te):
        return _rlib().og_log_k2it(C.byref(self.c), te)

    def b_planck(self, nu, te):
        return _rlib().og_b_planck(C.byref(self.c), nu, te)

    def gaunt(self, x, y):
        return _rlib().og_gaunt(C.byref(self.c), x, y)

    def j_ff(self, nu, te, ne, m_adm=M_ADM_DEFAULT):
        return _rlib().og_j_ff(C.byref(self.c), nu, te, ne, m_adm)

    def j_syn(self, nu, te, ne, b, cos_theta, m_adm=M_ADM_DEFAULT):
        return _rlib().og_j_syn(C.byref(self.c), nu, te, ne, b, cos_theta, m_adm)

    def thermal(self, nu, ne, te, b, bkcos, shift, I, tau, m_adm=M_ADM_DEFAULT):
        nu = _vec(nu)
        n = len(nu)
        I, tau = _vec(I), _vec(tau)
        dI, dtau = np.zeros(n), np.zeros(n)
        _rlib().og_rt_thermal(C.byref(self.c), n, _dp(nu), m_adm, ne, te, b, bkcos, shift,
                              _dp(I), _dp(tau), _dp(dI), _dp(dtau))
        return dI, dtau


_grid_set = set()   # variants whose argtypes are set


def _glib():
    L = lib()
    if _variant not in _grid_set:
        gp, dp = C.POINTER(Grid), C.POINTER(C.c_double)
        L.og_grid_down.argtypes = [gp, dp, dp, dp]
        L.og_grid_getuu.argtypes = [gp, dp, dp]
        L.og_grid_getuu.restype = C.c_double
        L.og_grid_rhs.argtypes = [gp, dp, dp, dp]
        L.og_grid_icond.argtypes = [gp] + [C.c_double] * 7 + [dp, dp]
        L.og_grid_getdt.argtypes = [gp, dp, dp, C.c_double]
        L.og_grid_getdt.restype = C.c_double
        L.og_grid_integrate.argtypes = [gp, dp, C.c_double]
        L.og_grid_evolve_ray.argtypes = [gp, dp, C.c_double, C.c_size_t, C.POINTER(C.c_int)]
        L.og_grid_evolve_ray.restype = C.c_int64
        L.og_grid_icond_all.argtypes = [gp, C.c_size_t, C.c_size_t] + [C.c_double] * 7 + [dp, dp]
        L.og_grid_evolve_all.argtypes = [gp, C.c_size_t, dp, dp, C.c_double, C.c_size_t,
                                         C.POINTER(C.c_int64), C.c_int]
        L.og_grid_evolve_all.restype = C.c_int64
        L.og_grid_sample.argtypes = [gp, C.c_int, C.c_int, dp]
        L.og_grid_sample.restype = C.c_double
        L.og_grid_rhs_all.argtypes = [gp, dp, dp]
        L.og_grid_flow.argtypes = [gp, dp, dp, dp, dp, dp]
        _grid_set.add(_variant)
    return L


class GridSpacetime:
    """Oracle view of a gridded spacetime: interleaved FP64 arrays
    gam[level] (n_points x 40) and met[level] (n_points x 10)."""

    def __init__(self, n, lo, hi, gam, met, times=(0.0, 0.0), map=MAP_UNIFORM, t_off=0.0,
                 raw_icond=False, ref_rule=False, rho=None, nu=(), m_adm=M_ADM_DEFAULT,
                 b_field=0.0):
        g = Grid()
        for a in range(3):
            g.n[a], g.lo[a], g.hi[a] = n[a], lo[a], hi[a]
        g.map = map
        g.t1, g.t2 = times
        g.t_off = t_off
        g.raw_icond, g.ref_rule = int(raw_icond), int(ref_rule)
        self._keep = []
        for l in range(2):
            ga = np.ascontiguousarray(gam[min(l, len(gam) - 1)], dtype=np.float64)
            me = np.ascontiguousarray(met[min(l, len(met) - 1)], dtype=np.float64)
            self._keep += [ga, me]
            g.gam[l], g.met[l] = ga.ctypes.data, me.ctypes.data
        if rho is not None:
            for l in range(2):
                r = np.ascontiguousarray(rho[min(l, len(rho) - 1)], dtype=np.float64)
                self._keep.append(r)
                g.rho[l] = r.ctypes.data
        elif len(nu):
            raise ValueError("transfer on a gridded spacetime needs the density (rho)")
        g.n_freq = len(nu)
        for i, v in enumerate(nu):
            g.nu[i] = v
        g.m_adm = m_adm
        g.b_field = b_field
        self.rt = Rt()
        g.rt = C.pointer(self.rt.c)
        self.n_data = 8 + 2 * len(nu)
        self.g = g

    def rhs_all(self, s):
        s = _vec(s)
        ds = np.zeros(self.n_data)
        _glib().og_grid_rhs_all(C.byref(self.g), _dp(s), _dp(ds))
        return ds

    def flow(self, q, u, with_bkcos=False):
        ne, shift, bkcos = C.c_double(), C.c_double(), C.c_double()
        _glib().og_grid_flow(C.byref(self.g), _dp(_vec(q)), _dp(_vec(u)), C.byref(ne),
                             C.byref(shift), C.byref(bkcos))
        if with_bkcos:
            return ne.value, shift.value, bkcos.value
        return ne.value, shift.value

    def sample(self, which, comp, q):
        return _glib().og_grid_sample(C.byref(self.g), which, comp, _dp(_vec(q)))

    def rhs(self, q, u):
        du = np.zeros(4)
        _glib().og_grid_rhs(C.byref(self.g), _dp(_vec(q)), _dp(_vec(u)), _dp(du))
        return du

    def down(self, q, u):
        k = np.zeros(4)
        _glib().og_grid_down(C.byref(self.g), _dp(_vec(q)), _dp(_vec(u)), _dp(k))
        return k

    def getuu(self, q, u):
        return _glib().og_grid_getuu(C.byref(self.g), _dp(_vec(q)), _dp(_vec(u)))

    def getdt(self, q, dt, u=(1.0, 0.0, 0.0, 0.0)):
        return _glib().og_grid_getdt(C.byref(self.g), _dp(_vec(q)), _dp(_vec(u)), dt)

    def icond(self, r_obs, i_obs, j_obs, alpha, beta):
        si, ci, sj, cj = camera_trig(i_obs, j_obs)
        q, u = np.zeros(4), np.zeros(4)
        _glib().og_grid_icond(C.byref(self.g), r_obs, si, ci, sj, cj, alpha, beta, _dp(q), _dp(u))
        return q, u

    def icond_all(self, w, h, w_img, h_img, r_obs, i_obs, j_obs):
        si, ci, sj, cj = camera_trig(i_obs, j_obs)
        data, info = np.zeros((w * h, self.n_data)), np.zeros(w * h)
        _glib().og_grid_icond_all(C.byref(self.g), w, h, w_img, h_img, r_obs, si, ci, sj, cj,
                                  _dp(data), _dp(info))
        return data, info

    def evolve_all(self, data, info, dt, n_sub, steps=None, n_threads=0):
        sp = steps.ctypes.data_as(C.POINTER(C.c_int64)) if steps is not None else None
        return _glib().og_grid_evolve_all(C.byref(self.g), data.shape[0], _dp(data), _dp(info),
                                          dt, n_sub, sp, n_threads)

    def render(self, w, h, w_img, h_img, r_obs, i_obs, j_obs, dt=-32.0, n_sub=1024, n_dump=1,
               n_threads=0):
        data, info = self.icond_all(w, h, w_img, h_img, r_obs, i_obs, j_obs)
        steps = np.zeros(w * h, dtype=np.int64)
        for _ in range(n_dump):
            self.evolve_all(data, info, dt, n_sub, steps, n_threads)
        return data, info, steps
