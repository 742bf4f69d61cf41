"""ctypes binding of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point as the
checker and by the benchmark's CPU-baseline leg.  The product path never
imports this module.
"""
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

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


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(_LIB)
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
        _lib = L
    return _lib


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
