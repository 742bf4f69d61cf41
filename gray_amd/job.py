"""ctypes binding of the gray job module libgray.so (include/gray_job.h) and
a reader of its raw dump files.

The dump layout is that of sim-org/io.c:39-72: four native size_t
(sz, n_data, w_rays, h_rays), the ray states in device order, then one
diagnostic real per ray.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgray.so")
LAUNCHER = os.path.join(_HERE, "lux-gray")

EXPORTS = ("gray_mkmod", "gray_rmmod", "gray_conf", "gray_init", "gray_exec", "gray_run",
           "gray_print_options", "gray_n_freq", "gray_n_data", "gray_get_option",
           "gray_write_raw", "gray_read_raw")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C gray_amd`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.gray_mkmod.argtypes = [vp]
        L.gray_mkmod.restype = vp
        L.gray_rmmod.argtypes = [vp]
        L.gray_rmmod.restype = None
        L.gray_conf.argtypes = [vp, C.c_char_p]
        L.gray_init.argtypes = [vp]
        L.gray_exec.argtypes = [vp]
        L.gray_run.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
        L.gray_n_freq.argtypes = [vp]
        L.gray_n_freq.restype = C.c_size_t
        L.gray_n_data.argtypes = [vp]
        L.gray_n_data.restype = C.c_size_t
        L.gray_get_option.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_size_t]
        L.gray_write_raw.argtypes = [C.c_char_p] + [C.c_size_t] * 4 + [vp, vp]
        L.gray_read_raw.argtypes = [C.c_char_p, vp, vp, vp]
        _lib = L
    return _lib


class Job:
    """mkmod / conf / init / exec / rmmod of one gray job."""

    def __init__(self):
        self._h = lib().gray_mkmod(None)
        if not self._h:
            raise MemoryError("gray_mkmod")

    def conf(self, arg):
        return lib().gray_conf(self._h, arg.encode())

    def init(self):
        return lib().gray_init(self._h)

    def exec(self):
        return lib().gray_exec(self._h)

    def get(self, name):
        buf = C.create_string_buffer(512)
        if lib().gray_get_option(self._h, name.encode(), buf, 512) != 0:
            raise KeyError(name)
        return buf.value.decode()

    @property
    def n_freq(self):
        return lib().gray_n_freq(self._h)

    @property
    def n_data(self):
        return lib().gray_n_data(self._h)

    def close(self):
        if self._h:
            lib().gray_rmmod(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def write_raw(path, data, info, w_rays, h_rays):
    data = np.ascontiguousarray(data)
    info = np.ascontiguousarray(info, dtype=data.dtype)
    sz = data.dtype.itemsize
    n_data = data.size // (w_rays * h_rays)
    rc = lib().gray_write_raw(path.encode(), sz, n_data, w_rays, h_rays,
                              data.ctypes.data, info.ctypes.data)
    if rc:
        raise OSError(f"gray_write_raw({path}) failed")


def load_raw(path):
    """(states [h, w, n_data], diagnostics [h, w]) of a raw dump in AoS
    order; SoA dumps come back as [n_data, h, w] via load_raw(..., soa)."""
    with open(path, "rb") as f:
        hdr = np.fromfile(f, dtype=np.uintp, count=4)
        sz, n, w, h = (int(x) for x in hdr)
        t = np.float64 if sz == 8 else np.float32
        states = np.fromfile(f, dtype=t, count=n * w * h).reshape(h, w, n)
        diag = np.fromfile(f, dtype=t, count=w * h).reshape(h, w)
    return states, diag
