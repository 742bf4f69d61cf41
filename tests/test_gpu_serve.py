"""Serving pool on the GPU: continuous batching of uneven requests.

* Each request's result is bit-for-bit the same alone and in a mix, with
  different pool sizes and epoch lengths (so its rays cross epochs at other
  steps and land in other lanes beside other requests' rays).
* Served Kerr images agree with the single-render FP32 form (image L1
  <= 1e-6) and with the FP64 oracle on a small image (L1 <= 1e-4); a served
  gridded request with transfer channels equals its single FP32-form render
  bit for bit and agrees with the FP64 grid oracle (image L1 <= 1e-4).
* A stratified sample of the seeded stream (every resolution up to 1024^2,
  analytic and gridded, 1-4 channels, two resident grids) is the same alone
  and in the mix, and each served image is within L1 1e-4 of the plain FP64
  form of the same request.
* Admission waits for HBM budget and for a free Kerr class; oversize
  requests and early result reads are errors.
"""
import numpy as np
import pytest

from gray_amd import grayhip as G
from gray_amd import spacetime as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NUS = (1.0, 2.67, 4.0, 8.0)


GNUS = (2.3e11, 5.0e11)


def bks_grid(n, rho_cap=100.0, dtype=4):
    """Boosted Kerr-Schild (a = 0.7, v = 0.3 along x) on n^3 over [-40, 40]^3,
    levels 0 and 1, with the generator's density 1/|x| capped at rho_cap (an
    odd n puts a grid point at the origin, where 1/|x| is ~1.8e308)."""
    xs = S.axis(-40, 40, n)
    b = S.bks(1.0, 0.7, 0.3, (1, 0, 0))
    st = G.Spacetime((n,) * 3, (-40.0,) * 3, (40.0,) * 3, dtype=dtype)
    tp = np.float32 if dtype == 4 else np.float64
    for l in range(2):
        g, m, r = S.bks_sample(b, float(l), xs, xs, xs, rho=True)
        st.upload(l, float(l), g.astype(tp), m.astype(tp))
        st.upload_rho(l, (r if rho_cap is None else np.minimum(r, rho_cap)).astype(tp))
    st.set_frozen(False)
    return st


@pytest.fixture(scope="module")
def grid():
    st = bks_grid(33)
    yield st
    st.close()


def mix(grid):
    return [
        G.kerr_request(24, 0.9375, 60.0, NUS),
        G.kerr_request(16, 0.3, 20.0, NUS[:1]),
        G.grid_request(16, grid, 45.0, nu=GNUS),
        G.kerr_request(32, 0.998, 85.0, NUS[:3], h_side=8),
        G.kerr_request(8, 0.0, 89.0, ()),
        G.grid_request(12, grid, 70.0, n_dump=3),
        G.kerr_request(20, 0.7, 40.0, NUS[:2], j_obs=30.0),
        G.grid_request(9, grid, 30.0, nu=GNUS[:1], h_side=13),
    ]


def image_l1(I, I0):
    """Relative L1 per channel over the rays finite in both images (the
    reference's thermal emission turns rays that end at the hole
    non-finite); also returns the share of those rays."""
    ok = np.isfinite(I).all(1) & np.isfinite(I0).all(1)
    den = np.abs(I0[ok]).sum(0)
    return np.abs(I[ok] - I0[ok]).sum(0) / np.where(den > 0, den, 1.0), ok.mean()


def run_all(reqs, **kw):
    with G.Server(**kw) as srv:
        ids = [srv.submit(q) for q in reqs]
        out = srv.run()
        steps, epochs = srv.stats()
    return [out[i] for i in ids], steps, epochs


def test_alone_equals_mixed(grid):
    reqs = mix(grid)
    mixed, steps, epochs = run_all(reqs, slots=1500, epoch_steps=700)
    assert epochs > 5 and steps[0] > 0 and steps[2] > 0
    for k, q in enumerate(mix(grid)):
        alone, _, _ = run_all([q], slots=333, epoch_steps=1234)
        for a, b in zip(alone[0], mixed[k]):
            assert np.array_equal(a, b, equal_nan=True), f"request {k} differs alone vs mixed"


def test_served_kerr_matches_render(grid):
    q = G.kerr_request(32, 0.9375, 60.0, NUS)
    (served,), _, _ = run_all([q], slots=4096, epoch_steps=2048)
    d, info, steps = served
    with G.Render(32, 32, 0.9375, precision=G.FP32, nu=NUS) as r:
        r.icond(64.0, 64.0, 1000.0, 60.0, 0.0)
        r.render(-32.0, 1024, 128)
        d1, _ = r.read_f64()
        s1 = r.steps()
    l1 = np.abs(d[:, 8:12] - d1[:, 8:12]).sum(0) / np.abs(d1[:, 8:12]).sum(0)
    assert (l1 <= 1e-6).all(), l1
    assert abs(steps.sum() / s1.sum() - 1) < 1e-3
    p = O.make_params(0.9375, nu=NUS)
    d0, _, _ = O.render(p, 32, 32, 64.0, 64.0, 1000.0, 60.0, 0.0)
    l1o = np.abs(d[:, 8:12] - d0[:, 8:12]).sum(0) / np.abs(d0[:, 8:12]).sum(0)
    assert (l1o <= 1e-4).all(), l1o


def test_served_channels_and_layout(grid):
    q2 = G.kerr_request(16, 0.5, 50.0, NUS[:2])
    q4 = G.kerr_request(16, 0.5, 50.0, NUS)
    (r2, r4), _, _ = run_all([q2, q4])
    d2, d4 = r2[0], r4[0]
    assert d2.shape == (256, 12) and d4.shape == (256, 16)
    # channel results do not depend on how many channels ride along
    np.testing.assert_array_equal(d2[:, :8], d4[:, :8])
    np.testing.assert_array_equal(d2[:, 8:10], d4[:, 8:10])   # I for nu = 1, 2.67
    np.testing.assert_array_equal(d2[:, 10:12], d4[:, 12:14])  # tau for nu = 1, 2.67


def test_admission_waits_for_budget():
    q = G.kerr_request(16, 0.5, 50.0, NUS)
    need = 256 * (16 + 2) * 8
    with G.Server(slots=64, hbm_budget=need + 256, epoch_steps=4096) as srv:
        a = srv.submit(q)
        b = srv.submit(G.kerr_request(16, 0.5, 50.0, NUS))
        srv.step()
        assert srv.state(a) == G.REQ_RUNNING and srv.state(b) == G.REQ_QUEUED
        with pytest.raises(G.GrayHipError):
            srv.result(a)  # not finished yet
        out = srv.run()
        assert set(out) == {a, b}
        np.testing.assert_array_equal(out[a][0], out[b][0])
        with pytest.raises(G.GrayHipError):
            srv.submit(G.kerr_request(64, 0.5, 50.0, NUS))  # larger than the budget


def test_submit_from_another_thread_while_stepping(grid):
    """One thread steps the pool while another submits the mix over time;
    every request finishes with the same result as in a single-thread run."""
    import threading
    import time
    reqs = mix(grid)
    ref, _, _ = run_all(reqs, slots=1500, epoch_steps=700)
    with G.Server(slots=1500, epoch_steps=700) as srv:
        ids = []

        def feed():
            for q in reqs:
                ids.append(srv.submit(q))
                time.sleep(0.002)

        th = threading.Thread(target=feed)
        th.start()
        out = {}
        while th.is_alive() or len(out) < len(reqs):
            srv.step()
            while True:
                rid = srv.next_done()
                if rid is None:
                    break
                out[rid] = srv.result(rid)
        th.join()
    assert len(out) == len(reqs)
    for k, rid in enumerate(ids):
        for a, b in zip(ref[k], out[rid]):
            assert np.array_equal(a, b, equal_nan=True), f"request {k} differs"


def test_served_grid_with_channels_matches_render_and_oracle(grid):
    """A gridded request with two channels (the reference flow on the
    grid's density): the served result agrees with the single FP32-form
    render (same step counts on >= 99% of rays, image L1 <= 1e-6; the
    serving kernel builds with multiply-add contraction, so its FP64 camera
    set-up may differ from the render's in the last bits) and with the FP64
    grid oracle (image L1 <= 1e-4)."""
    q = G.grid_request(16, grid, 45.0, nu=GNUS)
    (served,), _, _ = run_all([q], slots=700, epoch_steps=300)
    d, info, steps = served
    with G.Render(16, 16, precision=G.FP32, spacetime=grid, t_init=1.0, nu=GNUS) as r:
        r.icond(32.0, 32.0, 35.0, 45.0, 0.0)
        r.render(-32.0, 1024, 4)
        d1, i1 = r.read_f64()
        s1 = r.steps()
    assert (steps == s1).mean() >= 0.99
    l1r, fin = image_l1(d[:, 8:10], d1[:, 8:10])
    assert fin >= 0.95 and (l1r <= 1e-6).all(), (l1r, fin)
    n = 33
    xs = S.axis(-40, 40, n)
    b = S.bks(1.0, 0.7, 0.3, (1, 0, 0))
    lv = [S.bks_sample(b, float(l), xs, xs, xs, rho=True) for l in range(2)]
    og = O.GridSpacetime((n,) * 3, (-40.0,) * 3, (40.0,) * 3,
                         [g.astype(np.float32).astype(np.float64) for g, _, _ in lv],
                         [m.astype(np.float32).astype(np.float64) for _, m, _ in lv],
                         times=(0.0, 1.0), t_off=1.0,
                         rho=[np.minimum(r, 100.0).astype(np.float32).astype(np.float64)
                              for _, _, r in lv], nu=GNUS)
    d0, i0, s0 = og.render(16, 16, 32.0, 32.0, 35.0, 45.0, 0.0, n_dump=4)
    I0, I1 = d0[:, 8:10], d[:, 8:10]
    assert (np.abs(I0[np.isfinite(I0).all(1)]).sum(0) > 0).all()
    l1, fin = image_l1(I1, I0)
    assert fin >= 0.95 and (l1 <= 1e-4).all(), (l1, fin)


def test_admission_waits_for_a_free_kerr_class():
    """Nine requests whose schedules differ need nine Kerr classes: the
    ninth waits until one empties, and every result is the one it gets
    alone."""
    reqs = [G.kerr_request(8, 0.5, 50.0, NUS[:1], n_dump=8 + k) for k in range(9)]
    mixed, _, _ = run_all(reqs, slots=2048, epoch_steps=512)
    for k, q in enumerate(reqs):
        (alone,), _, _ = run_all([G.kerr_request(8, 0.5, 50.0, NUS[:1], n_dump=8 + k)],
                                 slots=300, epoch_steps=900)
        for a, b in zip(alone, mixed[k]):
            assert np.array_equal(a, b, equal_nan=True), k


def test_stratified_stream_sample_alone_mixed_and_fp64():
    """A stratified sample of the seeded request stream (seed 0x47524159):
    the first request of each resolution 64^2 .. 1024^2 (analytic or
    gridded as drawn) plus the first gridded requests on each of the two
    resident grids (96^3 and 128^3).  Each is the same bit for bit served in
    the mix and alone, and each served image is within L1 1e-4 of the plain
    FP64 form of the same request (gridded ones on FP64 copies of the
    grids)."""
    from gray_amd import serve as SV
    o = SV.defaults()
    items = SV.stream(o)
    pick, sides, grids_seen = [], set(), set()
    for it in items:
        if it.side not in sides:
            sides.add(it.side)
            pick.append(it)
        elif it.grid >= 0 and it.grid not in grids_seen and it.side <= 256:
            pick.append(it)
        if it.grid >= 0:
            grids_seen.add(it.grid)
        if len(sides) == 5 and grids_seen == {0, 1}:
            break
    assert {it.side for it in pick} == {64, 128, 256, 512, 1024}
    assert {it.grid for it in pick if it.grid >= 0} == {0, 1}
    sts = {}
    for k, n in enumerate((o.grid_n[0], o.grid_n[1])):
        sts[(k, 4)] = bks_grid(n, rho_cap=None)
        sts[(k, 8)] = bks_grid(n, rho_cap=None, dtype=8)

    def request(it):
        if it.grid >= 0:
            return G.grid_request(it.side, sts[(it.grid, 4)], it.i_obs,
                                  nu=SV.GRID_NUS[:it.n_freq])
        return G.kerr_request(it.side, it.a_spin, it.i_obs, SV.KERR_NUS[:it.n_freq])

    try:
        print(f"sample: {[(it.side, it.grid, it.n_freq) for it in pick]}", flush=True)
        mixed, steps, _ = run_all([request(it) for it in pick], slots=1 << 20)
        assert steps[0] > 0 and steps[2] > 0
        for k, it in enumerate(pick):
            alone, _, _ = run_all([request(it)], slots=1 << 18, epoch_steps=1500)
            for a, b in zip(alone[0], mixed[k]):
                assert np.array_equal(a, b, equal_nan=True), f"request {k} differs alone vs mixed"
            nf = it.n_freq
            if it.grid >= 0:
                r = G.Render(it.side, it.side, precision=G.FP64, spacetime=sts[(it.grid, 8)],
                             t_init=1.0, nu=SV.GRID_NUS[:nf])
                cam, n_dump = (32.0, 35.0), 4
            else:
                r = G.Render(it.side, it.side, it.a_spin, precision=G.FP64,
                             nu=SV.KERR_NUS[:nf])
                cam, n_dump = (64.0, 1000.0), 128
            with r:
                r.icond(cam[0], cam[0], cam[1], it.i_obs, 0.0)
                r.render(-32.0, 1024, n_dump)
                d64, _ = r.read_f64()
            I0, I1 = d64[:, 8:8 + nf], mixed[k][0][:, 8:8 + nf]
            l1, fin = image_l1(I1, I0)
            print(f"request {k}: {it.side}^2 grid {it.grid} nf {nf}: L1 vs FP64 {l1}, "
                  f"finite {fin:.4f}", flush=True)
            assert fin >= 0.95 and (l1 <= 1e-4).all(), (k, l1, fin)
    finally:
        for st in sts.values():
            st.close()
