"""Edge cases of the analytic-Kerr path on the GPU against the oracle:
degenerate and ragged image shapes, spin at 0, negative and near-extremal,
pole-on, edge-on and south-pole cameras, the maximum of 8 channels, a
zoomed plane entirely inside the shadow, and an 8192^2 image whose state
arrays pass 2^31 bytes.

Bars as in test_gpu_kerr.py: the plain FP64 form equals the oracle bit for
bit on geodesic state and step counts, transfer channels to 1e-10; the FP32
form stays within image relative L1 1e-4 of the oracle (checked where the
image has emission)."""
import numpy as np
import pytest

from gray_amd import grayhip as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NUS8 = (1.0, 1.5, 2.0, 2.67, 3.0, 4.0, 6.0, 8.0)


def both(a, w, h, w_img, h_img, i_obs, j_obs=0.0, nu=(), n_dump=128, precision=G.FP64):
    p = O.make_params(a, nu=nu)
    d0, i0, s0 = O.render(p, w, h, w_img, h_img, 1000.0, i_obs, j_obs, n_dump=n_dump)
    with G.Render(w, h, a, precision=precision, nu=nu) as r:
        r.icond(w_img, h_img, 1000.0, i_obs, j_obs)
        r.render(-32.0, 1024, n_dump)
        d1, i1 = r.read_f64()
        s1 = r.steps()
    return (d0, i0, s0), (d1, i1, s1)


def check_fp64(ref, got, nf):
    (d0, i0, s0), (d1, i1, s1) = ref, got
    assert np.array_equal(s0, s1), f"{(s0 != s1).sum()} rays differ in step count"
    assert np.array_equal(d0[:, :8], d1[:, :8])
    assert np.array_equal(i0, i1)
    if nf:
        rt0, rt1 = d0[:, 8:], d1[:, 8:]
        assert np.all(np.abs(rt1 - rt0) <= 1e-10 * np.abs(rt0)), np.abs(rt1 - rt0).max()


@pytest.mark.parametrize("w,h", [(1, 1), (7, 3), (3, 7), (1, 33)])
def test_shapes(w, h):
    check_fp64(*both(0.9375, w, h, 24.0, 16.0, 60.0, nu=(1.0, 4.0)), 2)


@pytest.mark.parametrize("a", [0.0, -0.7, 0.999])
def test_spins(a):
    check_fp64(*both(a, 12, 10, 24.0, 20.0, 60.0, j_obs=20.0, nu=(2.67,)), 1)


@pytest.mark.parametrize("a", [0.0, -0.7, 0.999])
def test_fp32_form_spins(a):
    """The FP32 form (pool, compensated state, FP64 re-integration below a
    5e-6 M step) against the oracle at the same spins, on a plane across
    the shadow's edge: for a = 0.999 the prograde photon orbit lies at
    r - r_H ~ 0.03, where the rays winding on it stay in the FP32 pool."""
    ref, got = both(a, 16, 16, 16.0, 16.0, 60.0, j_obs=20.0, nu=(2.67,), precision=G.FP32)
    (d0, _, s0), (d1, _, s1) = ref, got
    l1 = np.abs(d1[:, 8] - d0[:, 8]).sum() / np.abs(d0[:, 8]).sum()
    assert l1 <= 1e-4, l1
    assert abs(s1.sum() / s0.sum() - 1) < 5e-3, (s1.sum(), s0.sum())


@pytest.mark.parametrize("i_obs", [0.0, 90.0, 180.0])
def test_camera_orientations(i_obs):
    check_fp64(*both(0.9, 10, 10, 20.0, 20.0, i_obs, nu=(1.0,)), 1)


def test_eight_channels():
    ref, got = both(0.9375, 9, 9, 24.0, 24.0, 75.0, nu=NUS8)
    check_fp64(ref, got, 8)
    _, f32 = both(0.9375, 9, 9, 24.0, 24.0, 75.0, nu=NUS8, precision=G.FP32)
    I0, I1 = ref[0][:, 8:16], f32[0][:, 8:16]
    l1 = np.abs(I1 - I0).sum(0) / np.abs(I0).sum(0)
    assert (l1 <= 1e-4).all(), l1


def test_plane_inside_the_shadow():
    """A 2 M x 2 M plane at the centre of the shadow: every ray is captured
    and stops at r < 1.01 r_H well before the last interval."""
    ref, got = both(0.5, 8, 8, 2.0, 2.0, 30.0, n_dump=64)
    check_fp64(ref, got, 0)
    r = np.sqrt(np.sum(got[0][:, 1:4] ** 2, axis=1))
    assert (r < 2.0).all()
    assert (got[2] < 64 * 1024).all()


def test_restart_keeps_shared_depths_and_rejects_others():
    """grayhip_write then render continues a dump exactly as an uninterrupted
    run (plain FP64 form); depths not in the proportion of the disk's
    absorption law are refused."""
    a, nu, w = 0.9375, (1.0, 2.67, 4.0), 12
    with G.Render(w, w, a, precision=G.FP64, nu=nu) as r:
        r.icond(48.0, 48.0, 1000.0, 60.0, 0.0)
        r.render(-32.0, 1024, 40)
        mid, _ = r.read()
        r.render(-32.0, 1024, 88)
        ref, ref_info = r.read_f64()
    assert (mid[:, 11:14] > 0).any()
    with G.Render(w, w, a, precision=G.FP64, nu=nu) as r:
        r.write(mid)
        r.render(-32.0, 1024, 88)
        got, got_info = r.read_f64()
        assert np.array_equal(got, ref) and np.array_equal(got_info, ref_info)
        bad = mid.copy()
        k = int(np.argmax(bad[:, 11]))
        bad[k, 12] *= 1.001
        with pytest.raises(G.GrayHipError, match="optical depth"):
            r.write(bad)
    with G.Render(w, w, a, precision=G.FP32, nu=nu) as r:
        r.write(mid.astype(np.float32))
        r.render(-32.0, 1024, 88)
        got32, _ = r.read_f64()
    I0, I1 = ref[:, 8:11], got32[:, 8:11]
    l1 = np.abs(I1 - I0).sum(0) / np.abs(I0).sum(0)
    assert (l1 <= 1e-4).all(), l1


def test_fp64_lanes_outside_the_division_range():
    """The plain FP64 form's RHS and step rule take their correctly rounded
    divisions and square roots without the range steps only where a per-lane
    test puts every operand in range; lanes with a^2 z^2 subnormal or below
    2^-400 (z = 1e-160, 1e-100 near the equatorial plane), the discriminant
    (r^2 - a^2)^2 / 4 + a^2 z^2 beyond 2^600 (r = 1e60), and z = 0 (a^2 z^2 =
    0, in range) share waves with ordinary rays, and every lane stays bitwise
    equal to the oracle."""
    a, w = 0.9375, 8
    p = O.make_params(a)
    data, info = O.icond_all(p, w, w, 24.0, 24.0, 1000.0, 60.0, 0.0)
    for k, z in ((0, 1e-160), (9, 1e-100)):  # a^2 z^2 subnormal / below 2^-400
        data[k, 3], data[k, 7] = z, 0.0     # u^z = 0: z stays that small
    data[18, 3], data[18, 7] = 0.0, 0.0     # exactly in the plane
    for k in (27, 36):                      # r = 1e60, moving outward (dt < 0: inward)
        e = data[k, 1:4] / np.linalg.norm(data[k, 1:4])
        data[k, 1:4] = 1e60 * e
        data[k, 5:8] = np.linalg.norm(data[k, 5:8]) * e
    steps = np.zeros(w * w, dtype=np.int64)
    with G.Render(w, w, a, precision=G.FP64) as r:
        r.write(data)
        for _ in range(3):
            O.evolve_all(p, data, info, -32.0, 1024, steps)
        r.render(-32.0, 1024, 3)
        d1, i1 = r.read_f64()
        s1 = r.steps()
    assert (steps[[0, 9, 18, 27, 36]] > 0).all()
    assert 0 < abs(d1[0, 3]) < 1e-155 and 0 < abs(d1[9, 3]) < 1e-95 and d1[18, 3] == 0.0
    assert np.abs(d1[27, 1:4]).max() > 1e59
    assert np.array_equal(steps, s1), f"{(steps != s1).sum()} rays differ in step count"
    assert np.array_equal(d1, data) and np.array_equal(i1, info)


def test_large_image_8192_squared():
    """67,108,864 rays: the SoA state (8 reals a ray) spans 4.3 GB, so
    device offsets pass 2^31 bytes.  Two intervals of 16 sub-steps (the
    rays stay in the far field).  Plain FP64 form bit for bit the oracle on
    a lattice of pixels up to the last ray; FP32 form within 1e-6 relative
    of it on positions, with the same step counts."""
    w = h = 8192
    a = 0.9375
    cols = [0, 1, 2047, 4096, 6143, 8190, 8191]
    pix = [(gi, gj) for gj in cols for gi in cols]
    idx = np.array([gj * w + gi for gi, gj in pix])
    p = O.make_params(a)
    ref = O.icond_pixels(p, w, h, 64.0, 64.0, 1000.0, 60.0, 0.0, pix)
    ref_steps = O.evolve_subset(p, ref, 1, -32.0, 16, 2)
    out = {}
    for prec in (G.FP64, G.FP32):
        with G.Render(w, h, a, precision=prec) as r:
            r.icond(64.0, 64.0, 1000.0, 60.0, 0.0)
            r.render(-32.0, 16, 2)
            d, _ = r.read_f64()
            out[prec] = (d[idx].copy(), r.steps()[idx].copy(), int(r.steps().sum()))
            del d
    d64, s64, tot64 = out[G.FP64]
    d32, s32, tot32 = out[G.FP32]
    assert s64.sum() == ref_steps and tot64 > 16 * w * h
    assert np.array_equal(d64[:, :8], ref[:, :8])
    assert np.array_equal(s32, s64) and tot32 == tot64
    rel = np.abs(d32[:, 1:4] - d64[:, 1:4]).max(axis=1) / np.abs(d64[:, 1:4]).max(axis=1)
    assert rel.max() <= 1e-6, rel.max()
