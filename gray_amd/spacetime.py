"""ctypes binding of the numerical-spacetime tools in libgray.so
(include/gray_spacetime.h): the boosted Kerr-Schild sampler and HDF5 files
in the reference's layout."""
import ctypes as C

import numpy as np

from . import job as _job

EXPORTS = ("gray_bks_setup", "gray_bks_point", "gray_bks_sample", "gray_axis_uniform",
           "gray_axis_fisheye", "gray_h5_error", "gray_h5_create", "gray_h5_write_level",
           "gray_h5_list_times", "gray_h5_read_axes", "gray_h5_read_level", "gray_axis_map",
           "gray_h5_write_rays", "gray_h5_read_rays")


class BKS(C.Structure):
    _fields_ = [("m", C.c_double), ("a", C.c_double), ("v", C.c_double), ("n", C.c_double * 3)]


_set = False


def lib():
    global _set
    L = _job.lib()
    if not _set:
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        L.gray_bks_setup.argtypes = [C.POINTER(BKS), C.c_double, C.c_double, C.c_double, dp]
        L.gray_bks_setup.restype = None
        L.gray_bks_point.argtypes = [C.POINTER(BKS), dp, dp, dp]
        L.gray_bks_point.restype = None
        L.gray_bks_sample.argtypes = [C.POINTER(BKS), C.c_double, C.c_int, C.c_int, C.c_int,
                                      dp, dp, dp, vp, vp, vp]
        L.gray_bks_sample.restype = None
        L.gray_axis_uniform.argtypes = [C.c_double, C.c_double, C.c_int, dp]
        L.gray_axis_uniform.restype = None
        L.gray_axis_fisheye.argtypes = [C.c_double, C.c_double, C.c_int, dp]
        L.gray_axis_fisheye.restype = None
        L.gray_h5_error.restype = C.c_char_p
        L.gray_h5_create.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, dp, dp, dp, C.c_int]
        L.gray_h5_write_level.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int,
                                          vp, vp, vp, C.c_int]
        L.gray_h5_list_times.argtypes = [C.c_char_p, vp, dp, C.c_int]
        L.gray_h5_read_axes.argtypes = [C.c_char_p, C.POINTER(C.c_int), vp, vp, vp,
                                        C.POINTER(C.c_int)]
        L.gray_h5_read_level.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_int), vp, vp, vp,
                                         C.c_int]
        L.gray_axis_map.argtypes = [dp, C.c_int]
        L.gray_h5_write_rays.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, vp, C.c_int]
        L.gray_h5_read_rays.argtypes = [C.c_char_p, C.POINTER(C.c_int), vp]
        _set = True
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _vp(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def bks(m=1.0, a=0.0, v=0.0, axis=(1.0, 0.0, 0.0)):
    b = BKS()
    ax = np.asarray(axis, dtype=np.float64)
    lib().gray_bks_setup(C.byref(b), m, a, v, _dp(ax))
    return b


def bks_point(b, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    met, gam = np.zeros(10), np.zeros(40)
    lib().gray_bks_point(C.byref(b), _dp(x), _dp(met), _dp(gam))
    return met, gam


def axis(lo, hi, n, fisheye=False):
    out = np.zeros(n)
    (lib().gray_axis_fisheye if fisheye else lib().gray_axis_uniform)(lo, hi, n, _dp(out))
    return out


def axis_map(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().gray_axis_map(_dp(x), len(x))


def bks_sample(b, t, xs, ys, zs, rho=False):
    """Interleaved FP64 arrays gam [n, 40], met [n, 10] (and rho [n])."""
    xs, ys, zs = (np.ascontiguousarray(v, dtype=np.float64) for v in (xs, ys, zs))
    n = len(xs) * len(ys) * len(zs)
    gam, met = np.empty((n, 40)), np.empty((n, 10))
    r = np.empty(n) if rho else None
    lib().gray_bks_sample(C.byref(b), t, len(xs), len(ys), len(zs), _dp(xs), _dp(ys), _dp(zs),
                          _vp(gam), _vp(met), _vp(r))
    return (gam, met, r) if rho else (gam, met)


class H5Error(OSError):
    pass


def _h5(rc):
    if rc < 0:
        raise H5Error(lib().gray_h5_error().decode())
    return rc


def h5_create(path, xs, ys, zs, dtype=4):
    xs, ys, zs = (np.ascontiguousarray(v, dtype=np.float64) for v in (xs, ys, zs))
    _h5(lib().gray_h5_create(path.encode(), len(xs), len(ys), len(zs), _dp(xs), _dp(ys), _dp(zs),
                             dtype))


def h5_write_level(path, name, n, gam, met, rho=None, dtype=4):
    gam = np.ascontiguousarray(gam, dtype=np.float64)
    met = np.ascontiguousarray(met, dtype=np.float64)
    rho = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
    _h5(lib().gray_h5_write_level(path.encode(), name.encode(), n[0], n[1], n[2], _vp(gam),
                                  _vp(met), _vp(rho), dtype))


def h5_list_times(path, max_levels=1024):
    names = (C.c_char * 64 * max_levels)()
    times = np.zeros(max_levels)
    k = _h5(lib().gray_h5_list_times(path.encode(), names, _dp(times), max_levels))
    return [bytes(names[i]).split(b"\0")[0].decode() for i in range(k)], times[:k].copy()


def h5_read_axes(path):
    n = (C.c_int * 3)()
    dt = C.c_int()
    _h5(lib().gray_h5_read_axes(path.encode(), n, None, None, None, C.byref(dt)))
    xs, ys, zs = np.zeros(n[0]), np.zeros(n[1]), np.zeros(n[2])
    _h5(lib().gray_h5_read_axes(path.encode(), n, _vp(xs), _vp(ys), _vp(zs), C.byref(dt)))
    return xs, ys, zs, dt.value


def h5_read_level(path, name, n, dtype=8, rho=False):
    t = np.float32 if dtype == 4 else np.float64
    npnt = n[0] * n[1] * n[2]
    gam, met = np.empty((npnt, 40), dtype=t), np.empty((npnt, 10), dtype=t)
    r = np.empty(npnt, dtype=t) if rho else None
    nn = (C.c_int * 3)(*n)
    _h5(lib().gray_h5_read_level(path.encode(), name.encode(), nn, _vp(gam), _vp(met), _vp(r),
                                 dtype))
    return (gam, met, r) if rho else (gam, met)


def h5_write_rays(path, states, dtype=8):
    """states: [H, W, nd] in dump order; file /rays is [W, H, nd]."""
    s = np.ascontiguousarray(states, dtype=np.float64)
    h, w, nd = s.shape
    _h5(lib().gray_h5_write_rays(path.encode(), w, h, nd, _vp(s), dtype))


def h5_read_rays(path):
    """/rays of a ray file as FP64, in file order [W, H, nd]."""
    dims = (C.c_int * 3)()
    _h5(lib().gray_h5_read_rays(path.encode(), dims, None))
    out = np.empty(tuple(dims), dtype=np.float64)
    _h5(lib().gray_h5_read_rays(path.encode(), dims, _vp(out)))
    return out
