"""ctypes binding of the C request loop (include/gray_serve.h, libgray.so)."""
import ctypes as C

import numpy as np

from . import job as _job

EXPORTS = ("gray_serve_defaults", "gray_serve_run", "gray_serve_stream", "gray_serve_run_items")


class Opts(C.Structure):
    _fields_ = [
        ("n_requests", C.c_int),
        ("seed", C.c_uint64),
        ("rate", C.c_double),
        ("grid_fraction", C.c_double),
        ("max_side", C.c_int),
        ("grid_n", C.c_int * 2),
        ("slots", C.c_size_t),
        ("hbm_budget", C.c_size_t),
        ("epoch_steps", C.c_int),
        ("device", C.c_int),
        ("verbose", C.c_int),
    ]


class Report(C.Structure):
    _fields_ = [
        ("n_done", C.c_int),
        ("wall_s", C.c_double),
        ("p50_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("max_ms", C.c_double),
        ("mean_ms", C.c_double),
        ("ray_steps", C.c_int64 * 3),
        ("rays", C.c_int64),
        ("epochs", C.c_int64),
        ("ray_steps_per_s", C.c_double),
    ]


class Item(C.Structure):
    _fields_ = [
        ("side", C.c_int),
        ("a_spin", C.c_double),
        ("i_obs", C.c_double),
        ("n_freq", C.c_int),
        ("grid", C.c_int),
        ("arrival_s", C.c_double),
    ]


_set = False

# Gridded requests carry their drawn channels (the reference's thermal
# flow on the grid's density) as analytic ones do.
GRID_CHANNELS = True
# Channel frequencies of the request loop (gray_serve.c fill_request).
KERR_NUS = (1.0, 2.67, 4.0, 8.0)
GRID_NUS = (2.3e11, 3.45e11, 5.0e11, 8.6e11)


def served_channels(item):
    """Channels a served request of the stream carries."""
    return item.n_freq if (item.grid < 0 or GRID_CHANNELS) else 0


def lib():
    global _set
    L = _job.lib()
    if not _set:
        L.gray_serve_defaults.argtypes = [C.POINTER(Opts)]
        L.gray_serve_defaults.restype = None
        L.gray_serve_run.argtypes = [C.POINTER(Opts), C.POINTER(Report), C.c_void_p]
        L.gray_serve_stream.argtypes = [C.POINTER(Opts), C.c_void_p]
        L.gray_serve_run_items.argtypes = [C.POINTER(Opts), C.c_void_p, C.c_int,
                                           C.POINTER(Report), C.c_void_p, C.c_void_p]
        _set = True
    return L


def defaults(**kw):
    o = Opts()
    lib().gray_serve_defaults(C.byref(o))
    for k, v in kw.items():
        if k == "grid_n":
            o.grid_n[0], o.grid_n[1] = v
        else:
            setattr(o, k, v)
    return o


def stream(opts):
    items = (Item * opts.n_requests)()
    lib().gray_serve_stream(C.byref(opts), C.cast(items, C.c_void_p))
    return list(items)


def run(opts):
    """Run the request stream; returns (report dict, per-request latency ms)."""
    rep = Report()
    lat = np.zeros(opts.n_requests)
    rc = lib().gray_serve_run(C.byref(opts), C.byref(rep), C.c_void_p(lat.ctypes.data))
    if rc != 0:
        raise RuntimeError("gray_serve_run failed")
    d = {f: getattr(rep, f) for f, _ in Report._fields_}
    d["ray_steps"] = list(rep.ray_steps)
    return d, lat


def run_items(opts, items, keep_results=False):
    """Serve the given items (a list of Item, arrivals in time order).
    Returns (report dict, per-request latency ms, results or None), results
    being each request's AoS FP64 rays [side*side, 8 + 2 n_freq]."""
    n = len(items)
    arr = (Item * n)(*items)
    rep = Report()
    lat = np.zeros(n)
    res = None
    ptrs = None
    if keep_results:
        res = [np.zeros((it.side * it.side, 8 + 2 * served_channels(it))) for it in items]
        ptrs = (C.c_void_p * n)(*[r.ctypes.data for r in res])
    rc = lib().gray_serve_run_items(C.byref(opts), C.cast(arr, C.c_void_p), n, C.byref(rep),
                                    C.c_void_p(lat.ctypes.data),
                                    C.cast(ptrs, C.c_void_p) if ptrs is not None else None)
    if rc != 0:
        raise RuntimeError("gray_serve_run_items failed")
    d = {f: getattr(rep, f) for f, _ in Report._fields_}
    d["ray_steps"] = list(rep.ray_steps)
    return d, lat, res
