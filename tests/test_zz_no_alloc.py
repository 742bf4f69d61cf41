"""grayhip_render_static (the render that allocates nothing) against
grayhip_render through the C ABI: two calls in a row on one stream, the
second smaller, each bit for bit equal to a context render of the same
job (state, g(u,u), steps); sizes beyond the static scratch are refused.
Device buffers come from the HIP runtime the library itself uses."""
import ctypes as C

import numpy as np
import pytest

from gray_amd import grayhip as G

pytestmark = pytest.mark.gpu

NUS = (1.0, 2.67, 4.0)


def hip():
    G.lib()                      # loads the HIP runtime libgrayhip.so links
    h = C.CDLL("libamdhip64.so.7")
    vp = C.c_void_p
    h.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
    h.hipFree.argtypes = [vp]
    h.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(vp)]
    h.hipStreamSynchronize.argtypes = [vp]
    h.hipStreamDestroy.argtypes = [vp]
    return h


D2H = 2


def config(w, h, a, precision, nu):
    cfg = G.Config()
    cfg.w_rays, cfg.h_rays, cfg.precision, cfg.a_spin = w, h, precision, a
    cfg.n_freq = len(nu)
    for i, v in enumerate(nu):
        cfg.nu[i] = v
    cfg.fluid = G.FLUID_DISK if nu else G.FLUID_NONE
    cfg.escape, cfg.r_esc = G.ESC_OUTWARD, 1000.0
    return cfg


def render_static(H, stream, w, h, a, i_obs, precision, nu, n_dump):
    n, nd = w * h, 8 + 2 * len(nu)
    bufs = [C.c_void_p() for _ in range(3)]
    sizes = (n * nd * 8, n * 8, n * 8)
    for b, s in zip(bufs, sizes):
        assert H.hipMalloc(C.byref(b), s) == 0
    try:
        cfg = config(w, h, a, precision, nu)
        rc = G.lib().grayhip_render_static(C.byref(cfg), 64.0, 64.0, 1000.0, i_obs, 0.0, -32.0,
                                           1024, n_dump, bufs[0], bufs[1], bufs[2], stream)
        assert rc == 0, G.lib().grayhip_global_error().decode()
        assert H.hipStreamSynchronize(stream) == 0
        soa = np.empty((nd, n))
        info = np.empty(n)
        steps = np.empty(n, dtype=np.int64)
        for arr, b in zip((soa, info, steps), bufs):
            assert H.hipMemcpy(arr.ctypes.data, b, arr.nbytes, D2H) == 0
    finally:
        for b in bufs:
            H.hipFree(b)
    return soa.T.copy(), info, steps


def render_ctx(w, h, a, i_obs, precision, nu, n_dump):
    with G.Render(w, h, a, precision=precision, nu=nu) as r:
        r.icond(64.0, 64.0, 1000.0, i_obs, 0.0)
        r.render(-32.0, 1024, n_dump)
        d, info = r.read_f64()
        return d, info, r.steps()


@pytest.mark.parametrize("precision", [G.FP32, G.FP64])
def test_static_render_matches_context_render_twice(precision):
    H = hip()
    stream = C.c_void_p()
    assert H.hipStreamCreate(C.byref(stream)) == 0
    try:
        # a = 0.998 near edge-on: rays hand over to the FP64 re-integration
        for (w, h, a, i_obs, nu) in ((96, 80, 0.998, 85.0, NUS), (40, 24, 0.9375, 60.0, NUS[:1])):
            print(f"static render {w}x{h} a={a} i={i_obs} nf={len(nu)} precision={precision}",
                  flush=True)
            got = render_static(H, stream, w, h, a, i_obs, precision, nu, 128)
            want = render_ctx(w, h, a, i_obs, precision, nu, 128)
            for x, y in zip(got, want):
                assert np.array_equal(x, y)
            assert got[2].sum() > 0
    finally:
        H.hipStreamDestroy(stream)


def test_static_render_refuses_oversize():
    cfg = config(1025, 1024, 0.5, G.FP32, ())
    dummy = C.c_void_p(1)
    rc = G.lib().grayhip_render_static(C.byref(cfg), 64.0, 64.0, 1000.0, 60.0, 0.0, -32.0, 1024,
                                       8, dummy, dummy, dummy, None)
    assert rc == -1
    assert "static scratch" in G.lib().grayhip_global_error().decode()
