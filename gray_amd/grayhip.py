"""ctypes binding of the gfx950 device layer libgrayhip.so (include/grayhip.h).

The library is built in-tree by ``make -C gray_amd``; importing this module
without it raises, so nothing silently falls back to CPU code.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgrayhip.so")

MAX_FREQ = 8
FP32, FP64 = 4, 8
ESC_REF, ESC_OUTWARD = 0, 1
FLUID_NONE, FLUID_DISK, FLUID_UNIFORM, FLUID_THERMAL = 0, 1, 2, 3

# Every entry point declared in include/grayhip.h.
EXPORTS = (
    "grayhip_create", "grayhip_destroy", "grayhip_icond", "grayhip_evolve",
    "grayhip_render", "grayhip_read", "grayhip_read_f64", "grayhip_write",
    "grayhip_read_steps", "grayhip_stats", "grayhip_pass_steps", "grayhip_kernel_time",
    "grayhip_n_data", "grayhip_read_bl", "grayhip_read_covariant", "grayhip_render_static",
    "grayhip_last_error", "grayhip_global_error", "grayhip_device_count",
    "grayhip_spacetime_create", "grayhip_spacetime_destroy", "grayhip_spacetime_upload",
    "grayhip_spacetime_shift", "grayhip_spacetime_set_frozen", "grayhip_spacetime_bytes",
    "grayhip_spacetime_sample", "grayhip_spacetime_error", "grayhip_spacetime_upload_rho",
)
MAP_UNIFORM, MAP_FISHEYE = 0, 1


class GridDesc(C.Structure):
    _fields_ = [
        ("n", C.c_int * 3),
        ("lo", C.c_double * 3),
        ("hi", C.c_double * 3),
        ("map", C.c_int),
        ("dtype", C.c_int),
        ("device", C.c_int),
    ]


class Config(C.Structure):
    _fields_ = [
        ("w_rays", C.c_size_t),
        ("h_rays", C.c_size_t),
        ("precision", C.c_int),
        ("a_spin", C.c_double),
        ("n_freq", C.c_int),
        ("nu", C.c_double * MAX_FREQ),
        ("fluid", C.c_int),
        ("escape", C.c_int),
        ("r_esc", C.c_double),
        ("eps_stop", C.c_double),
        ("device", C.c_int),
        ("spacetime", C.c_void_p),
        ("t_init", C.c_double),
        ("raw_icond", C.c_int),
        ("m_adm", C.c_double),
        ("b_field", C.c_double),
    ]


class GrayHipError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libgrayhip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GrayHipError(f"{LIB_PATH} is missing: build it with `make -C gray_amd`")
    L = C.CDLL(LIB_PATH)
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    L.grayhip_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
    L.grayhip_destroy.argtypes = [vp]
    L.grayhip_destroy.restype = None
    L.grayhip_icond.argtypes = [vp] + [C.c_double] * 5
    L.grayhip_evolve.argtypes = [vp, C.c_double, C.c_size_t, dp]
    L.grayhip_render.argtypes = [vp, C.c_double, C.c_size_t, C.c_size_t, dp]
    L.grayhip_read.argtypes = [vp, vp, vp]
    L.grayhip_read_f64.argtypes = [vp, vp, vp]
    L.grayhip_write.argtypes = [vp, vp]
    L.grayhip_read_steps.argtypes = [vp, vp]
    L.grayhip_stats.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.grayhip_pass_steps.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.grayhip_kernel_time.argtypes = [vp, dp, C.POINTER(C.c_int)]
    L.grayhip_n_data.argtypes = [vp]
    L.grayhip_read_bl.argtypes = [vp, C.c_double, vp]
    L.grayhip_read_covariant.argtypes = [vp, vp]
    L.grayhip_render_static.argtypes = [C.POINTER(Config)] + [C.c_double] * 6 + \
        [C.c_size_t, C.c_size_t, vp, vp, vp, vp]
    L.grayhip_last_error.argtypes = [vp]
    L.grayhip_last_error.restype = C.c_char_p
    L.grayhip_global_error.argtypes = []
    L.grayhip_global_error.restype = C.c_char_p
    L.grayhip_device_count.argtypes = []
    L.grayhip_spacetime_create.argtypes = [C.POINTER(GridDesc), C.POINTER(vp)]
    L.grayhip_spacetime_destroy.argtypes = [vp]
    L.grayhip_spacetime_destroy.restype = None
    L.grayhip_spacetime_upload.argtypes = [vp, C.c_int, C.c_double, vp, vp]
    L.grayhip_spacetime_shift.argtypes = [vp, C.c_int]
    L.grayhip_spacetime_set_frozen.argtypes = [vp, C.c_int]
    L.grayhip_spacetime_bytes.argtypes = [vp]
    L.grayhip_spacetime_bytes.restype = C.c_size_t
    L.grayhip_spacetime_sample.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int, C.c_double, vp]
    L.grayhip_spacetime_upload_rho.argtypes = [vp, C.c_int, vp]
    L.grayhip_spacetime_error.argtypes = [vp]
    L.grayhip_spacetime_error.restype = C.c_char_p
    _lib = L
    return L


def device_count():
    return lib().grayhip_device_count()


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class Spacetime:
    """A numerical spacetime resident on the GPU (two time-level slots)."""

    def __init__(self, n, lo, hi, map=MAP_UNIFORM, dtype=4, device=0):
        d = GridDesc()
        for a in range(3):
            d.n[a], d.lo[a], d.hi[a] = n[a], lo[a], hi[a]
        d.map, d.dtype, d.device = map, dtype, device
        self.desc = d
        self.n_points = n[0] * n[1] * n[2]
        self.np_dtype = np.float32 if dtype == 4 else np.float64
        self._h = C.c_void_p()
        rc = lib().grayhip_spacetime_create(C.byref(d), C.byref(self._h))
        if rc != 0:
            raise GrayHipError(f"grayhip_spacetime_create: {rc}")

    def _check(self, rc, what):
        if rc != 0:
            raise GrayHipError(f"{what}: {rc}: {lib().grayhip_spacetime_error(self._h).decode()}")

    def upload(self, slot, time, gam, met):
        gam = np.ascontiguousarray(gam, dtype=self.np_dtype)
        met = np.ascontiguousarray(met, dtype=self.np_dtype)
        assert gam.size == 40 * self.n_points and met.size == 10 * self.n_points
        self._check(lib().grayhip_spacetime_upload(self._h, slot, time, _ptr(gam), _ptr(met)),
                    "upload")

    def upload_rho(self, slot, rho):
        rho = np.ascontiguousarray(rho, dtype=self.np_dtype)
        assert rho.size == self.n_points
        self._check(lib().grayhip_spacetime_upload_rho(self._h, slot, _ptr(rho)), "upload_rho")

    def shift(self, to_t1=True):
        self._check(lib().grayhip_spacetime_shift(self._h, int(to_t1)), "shift")

    def set_frozen(self, frozen):
        self._check(lib().grayhip_spacetime_set_frozen(self._h, int(frozen)), "set_frozen")

    def bytes(self):
        return lib().grayhip_spacetime_bytes(self._h)

    def sample(self, events, which, comp, t_init=0.0):
        ev = np.ascontiguousarray(events, dtype=np.float64).reshape(-1, 4)
        out = np.empty(len(ev))
        self._check(lib().grayhip_spacetime_sample(self._h, len(ev), _ptr(ev), which, comp,
                                                   t_init, _ptr(out)), "sample")
        return out

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().grayhip_spacetime_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Render:
    """One render on the GPU: W*H rays of one camera, one spacetime."""

    def __init__(self, w_rays, h_rays, a_spin=0.0, precision=FP64, nu=(), fluid=None,
                 escape=ESC_OUTWARD, r_esc=1000.0, eps_stop=0.0, device=0,
                 spacetime=None, t_init=0.0, raw_icond=False, m_adm=0.0, b_field=0.0):
        cfg = Config()
        cfg.w_rays, cfg.h_rays = w_rays, h_rays
        cfg.precision = precision
        cfg.a_spin = a_spin
        cfg.n_freq = len(nu)
        for i, v in enumerate(nu):
            cfg.nu[i] = v
        if fluid is None:
            fluid = (FLUID_THERMAL if spacetime is not None else FLUID_DISK) if len(nu) \
                else FLUID_NONE
        cfg.fluid = fluid
        cfg.escape = escape
        cfg.r_esc = r_esc
        cfg.eps_stop = eps_stop
        cfg.device = device
        cfg.spacetime = spacetime.handle if spacetime is not None else None
        cfg.t_init = t_init
        cfg.raw_icond = int(raw_icond)
        cfg.m_adm = m_adm
        cfg.b_field = b_field
        self._spacetime = spacetime  # keep it alive while the render exists
        self.cfg = cfg
        self.n_rays = w_rays * h_rays
        self.n_data = 8 + 2 * len(nu)
        self._h = C.c_void_p()
        rc = lib().grayhip_create(C.byref(cfg), C.byref(self._h))
        if rc != 0:
            raise GrayHipError(f"grayhip_create: {rc}: {lib().grayhip_global_error().decode()}")

    def _check(self, rc, what):
        if rc != 0:
            raise GrayHipError(f"{what}: {rc}: {lib().grayhip_last_error(self._h).decode()}")

    def close(self):
        if self._h:
            lib().grayhip_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def icond(self, w_img, h_img, r_obs, i_obs, j_obs):
        self._check(lib().grayhip_icond(self._h, w_img, h_img, r_obs, i_obs, j_obs), "icond")

    def evolve(self, dt, n_sub):
        ns = C.c_double()
        self._check(lib().grayhip_evolve(self._h, dt, n_sub, C.byref(ns)), "evolve")
        return ns.value

    def render(self, dt, n_sub, n_dump):
        ns = C.c_double()
        self._check(lib().grayhip_render(self._h, dt, n_sub, n_dump, C.byref(ns)), "render")
        return ns.value

    def read(self):
        """(data AoS [n_rays, n_data], info [n_rays]) in the configured precision."""
        t = np.float64 if self.cfg.precision == FP64 else np.float32
        data = np.empty((self.n_rays, self.n_data), dtype=t)
        info = np.empty(self.n_rays, dtype=t)
        self._check(lib().grayhip_read(self._h, _ptr(data), _ptr(info)), "read")
        return data, info

    def read_f64(self):
        data = np.empty((self.n_rays, self.n_data), dtype=np.float64)
        info = np.empty(self.n_rays, dtype=np.float64)
        self._check(lib().grayhip_read_f64(self._h, _ptr(data), _ptr(info)), "read_f64")
        return data, info

    def read_bl(self, r_match):
        """Boyer-Lindquist states [n_rays, 8]: t, r, theta, phi, u^t, u^r, u^theta, u^phi."""
        out = np.empty((self.n_rays, 8), dtype=np.float64)
        self._check(lib().grayhip_read_bl(self._h, r_match, _ptr(out)), "read_bl")
        return out

    def read_covariant(self):
        """States [n_rays, 8]: t, x, y, z, k_t, k_x, k_y, k_z (k = g u)."""
        out = np.empty((self.n_rays, 8), dtype=np.float64)
        self._check(lib().grayhip_read_covariant(self._h, _ptr(out)), "read_covariant")
        return out

    def write(self, data):
        t = np.float64 if self.cfg.precision == FP64 else np.float32
        data = np.ascontiguousarray(data, dtype=t)
        assert data.shape == (self.n_rays, self.n_data)
        self._check(lib().grayhip_write(self._h, _ptr(data)), "write")

    def steps(self):
        s = np.empty(self.n_rays, dtype=np.int64)
        self._check(lib().grayhip_read_steps(self._h, _ptr(s)), "read_steps")
        return s

    def stats(self):
        st, fb = C.c_int64(), C.c_int64()
        self._check(lib().grayhip_stats(self._h, C.byref(st), C.byref(fb)), "stats")
        return st.value, fb.value

    def pass_steps(self):
        a, b = C.c_int64(), C.c_int64()
        self._check(lib().grayhip_pass_steps(self._h, C.byref(a), C.byref(b)), "pass_steps")
        return a.value, b.value

    def kernel_time(self):
        ns, n = C.c_double(), C.c_int()
        self._check(lib().grayhip_kernel_time(self._h, C.byref(ns), C.byref(n)), "kernel_time")
        return ns.value, n.value


# ---------------------------------------------------------------------------
# serving pool (grayhip_server_*)

REQ_QUEUED, REQ_RUNNING, REQ_DONE = 0, 1, 2
EXPORTS = EXPORTS + (
    "grayhip_server_create", "grayhip_server_destroy", "grayhip_server_submit",
    "grayhip_server_step", "grayhip_server_next_done", "grayhip_server_state",
    "grayhip_server_result", "grayhip_server_stats", "grayhip_server_error",
)


class ServerConfig(C.Structure):
    _fields_ = [
        ("device", C.c_int),
        ("slots", C.c_size_t),
        ("hbm_budget", C.c_size_t),
        ("epoch_steps", C.c_int),
        ("max_requests", C.c_int),
    ]


class Request(C.Structure):
    _fields_ = [
        ("w_rays", C.c_size_t),
        ("h_rays", C.c_size_t),
        ("a_spin", C.c_double),
        ("n_freq", C.c_int),
        ("nu", C.c_double * 4),
        ("fluid", C.c_int),
        ("escape", C.c_int),
        ("r_esc", C.c_double),
        ("eps_stop", C.c_double),
        ("w_img", C.c_double),
        ("h_img", C.c_double),
        ("r_obs", C.c_double),
        ("i_obs", C.c_double),
        ("j_obs", C.c_double),
        ("dt", C.c_double),
        ("n_sub", C.c_size_t),
        ("n_dump", C.c_size_t),
        ("spacetime", C.c_void_p),
        ("t_init", C.c_double),
    ]


_srv_set = False


def _slib():
    global _srv_set
    L = lib()
    if not _srv_set:
        vp = C.c_void_p
        L.grayhip_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(vp)]
        L.grayhip_server_destroy.argtypes = [vp]
        L.grayhip_server_destroy.restype = None
        L.grayhip_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_int64)]
        L.grayhip_server_step.argtypes = [vp, C.POINTER(C.c_int)]
        L.grayhip_server_next_done.argtypes = [vp, C.POINTER(C.c_int64)]
        L.grayhip_server_state.argtypes = [vp, C.c_int64, C.POINTER(C.c_int)]
        L.grayhip_server_result.argtypes = [vp, C.c_int64, vp, vp, vp]
        L.grayhip_server_stats.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.grayhip_server_error.argtypes = [vp]
        L.grayhip_server_error.restype = C.c_char_p
        _srv_set = True
    return L


def kerr_request(side, a_spin, i_obs, nu=(), j_obs=0.0, w_img=64.0, r_obs=1000.0, n_dump=128,
                 dt=-32.0, n_sub=1024, h_side=None):
    q = Request()
    q.w_rays, q.h_rays = side, h_side or side
    q.a_spin = a_spin
    q.n_freq = len(nu)
    for i, v in enumerate(nu):
        q.nu[i] = v
    q.fluid = FLUID_DISK if nu else FLUID_NONE
    q.escape = ESC_OUTWARD
    q.r_esc = 1000.0
    q.w_img = q.h_img = w_img
    q.r_obs, q.i_obs, q.j_obs = r_obs, i_obs, j_obs
    q.dt, q.n_sub, q.n_dump = dt, n_sub, n_dump
    return q


def grid_request(side, spacetime, i_obs, w_img=32.0, r_obs=35.0, n_dump=4, t_init=1.0,
                 dt=-32.0, n_sub=1024, nu=(), h_side=None):
    """A gridded request; channels nu use the reference's thermal flow on
    the spacetime's density (it must be uploaded)."""
    q = Request()
    q.w_rays, q.h_rays = side, h_side or side
    q.n_freq = len(nu)
    for i, v in enumerate(nu):
        q.nu[i] = v
    q.fluid = FLUID_THERMAL if nu else FLUID_NONE
    q.escape = ESC_OUTWARD
    q.w_img = q.h_img = w_img
    q.r_obs, q.i_obs = r_obs, i_obs
    q.dt, q.n_sub, q.n_dump = dt, n_sub, n_dump
    q.spacetime = spacetime.handle
    q.t_init = t_init
    return q


class Server:
    """The serving pool: submit requests, step epochs, collect results."""

    def __init__(self, slots=1 << 16, hbm_budget=1 << 30, epoch_steps=2048, max_requests=1024,
                 device=0):
        c = ServerConfig(device, slots, hbm_budget, epoch_steps, max_requests)
        self._h = C.c_void_p()
        rc = _slib().grayhip_server_create(C.byref(c), C.byref(self._h))
        if rc != 0:
            raise GrayHipError(f"grayhip_server_create: {rc}: "
                               f"{_slib().grayhip_server_error(None).decode()}")
        self._meta = {}

    def _check(self, rc, what):
        if rc != 0:
            raise GrayHipError(f"{what}: {rc}: {_slib().grayhip_server_error(self._h).decode()}")

    def submit(self, q):
        i = C.c_int64()
        self._check(_slib().grayhip_server_submit(self._h, C.byref(q), C.byref(i)), "submit")
        self._meta[i.value] = (q.w_rays * q.h_rays, 8 + 2 * q.n_freq)
        return i.value

    def step(self):
        busy = C.c_int()
        self._check(_slib().grayhip_server_step(self._h, C.byref(busy)), "step")
        return bool(busy.value)

    def next_done(self):
        i = C.c_int64()
        rc = _slib().grayhip_server_next_done(self._h, C.byref(i))
        if rc == 1:
            return None
        self._check(rc, "next_done")
        return i.value

    def state(self, rid):
        s = C.c_int()
        self._check(_slib().grayhip_server_state(self._h, rid, C.byref(s)), "state")
        return s.value

    def result(self, rid):
        n, nd = self._meta[rid]
        data, info = np.empty((n, nd)), np.empty(n)
        steps = np.empty(n, dtype=np.int64)
        self._check(_slib().grayhip_server_result(self._h, rid, _ptr(data), _ptr(info), _ptr(steps)),
                    "result")
        del self._meta[rid]
        return data, info, steps

    def stats(self):
        st = (C.c_int64 * 3)()
        ep = C.c_int64()
        self._check(_slib().grayhip_server_stats(self._h, st, C.byref(ep)), "stats")
        return list(st), ep.value

    def run(self, max_epochs=100000):
        """Step until idle; returns {id: (data, info, steps)} of all results."""
        out = {}
        for _ in range(max_epochs):
            busy = self.step()
            while True:
                rid = self.next_done()
                if rid is None:
                    break
                out[rid] = self.result(rid)
            if not busy:
                return out
        raise GrayHipError("server did not drain")

    def close(self):
        if self._h:
            _slib().grayhip_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
