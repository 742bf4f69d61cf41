"""Serving pool on the GPU: continuous batching of uneven requests.

* Each request's result is bit-for-bit the same alone and in a mix, with
  different pool sizes and epoch lengths (so its rays cross epochs at other
  steps and land in other slots beside other requests' rays).
* Served Kerr images agree with the single-render FP32 form (image L1
  <= 1e-6) and with the FP64 oracle on a small image (L1 <= 1e-4).
* Admission waits for HBM budget; oversize requests and early result
  reads are errors.
"""
import numpy as np
import pytest

from gray_amd import grayhip as G
from gray_amd import spacetime as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NUS = (1.0, 2.67, 4.0, 8.0)


@pytest.fixture(scope="module")
def grid():
    n = 33
    xs = S.axis(-40, 40, n)
    b = S.bks(1.0, 0.7, 0.3, (1, 0, 0))
    st = G.Spacetime((n,) * 3, (-40.0,) * 3, (40.0,) * 3, dtype=4)
    for l in range(2):
        g, m = S.bks_sample(b, float(l), xs, xs, xs)
        st.upload(l, float(l), g.astype(np.float32), m.astype(np.float32))
    st.set_frozen(False)
    yield st
    st.close()


def mix(grid):
    return [
        G.kerr_request(24, 0.9375, 60.0, NUS),
        G.kerr_request(16, 0.3, 20.0, NUS[:1]),
        G.grid_request(16, grid, 45.0),
        G.kerr_request(32, 0.998, 85.0, NUS[:3], h_side=8),
        G.kerr_request(8, 0.0, 89.0, ()),
        G.grid_request(12, grid, 70.0, n_dump=3),
        G.kerr_request(20, 0.7, 40.0, NUS[:2], j_obs=30.0),
    ]


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
            assert np.array_equal(a, b), f"request {k} differs alone vs mixed"


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
            assert np.array_equal(a, b), f"request {k} differs"
