"""The lux-gray launcher end to end on the GPU: dump files of a short
analytic-Kerr job against the oracle (plain FP64 form: bitwise)."""
import os
import subprocess

import numpy as np
import pytest

from gray_amd import job as J
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def run(tmp_path, *args):
    r = subprocess.run([J.LAUNCHER, "gray", *args, f"out={tmp_path}/%04d.raw"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_launcher_kerr_dumps_match_oracle(tmp_path):
    a, nus = 0.9375, (1.0, 4.0)
    out = run(tmp_path, "coor=KS", "a=0.9375", "W=24", "H=16", "w=48", "h=32", "r=1000",
              "i=60", "j=15", "p=8", "N=3", "nu=1,4")
    assert "ns/step/ray" in out and "3: -64.0 -> -96.0" in out
    p = O.make_params(a, nu=nus)
    data, info = O.icond_all(p, 24, 16, 48.0, 32.0, 1000.0, 60.0, 15.0)
    for i in range(4):
        s, d = J.load_raw(f"{tmp_path}/{i:04d}.raw")
        assert s.shape == (16, 24, 12) and s.dtype == np.float64
        assert np.array_equal(s.reshape(-1, 12)[:, :8], data[:, :8]), i
        rel = np.abs(s.reshape(-1, 12) - data) / np.maximum(np.abs(data), 1e-300)
        assert rel.max() <= 1e-10
        # g(u,u) of every dump: the oracle's info after the same intervals
        assert d.shape == (16, 24) and np.array_equal(d.ravel(), info), i
        O.evolve_all(p, data, info, -32.0, 1024)


def test_launcher_soa_and_single_precision(tmp_path):
    run(tmp_path, "coor=KS", "a=0.5", "W=8", "H=4", "p=4", "N=1", "mo=SoA")
    hdr = np.fromfile(f"{tmp_path}/0000.raw", dtype=np.uintp, count=4)
    assert list(hdr) == [4, 8, 8, 4]
    raw = np.fromfile(f"{tmp_path}/0000.raw", dtype=np.float32, offset=32)
    soa = raw[:8 * 32].reshape(8, 32)
    p = O.make_params(0.5)
    data, _ = O.icond_all(p, 8, 4, 64.0, 64.0, 1024.0, 60.0, 0.0)
    np.testing.assert_array_equal(soa.T, data.astype(np.float32))
    assert os.path.exists(f"{tmp_path}/0001.raw")


def test_rewrite_surface_ray_files_match_oracle(tmp_path):
    """spacetime=Kerr (sim/gray.c): infcam rays, n_dump intervals ending at
    dt_dump * (i + 1), ray%04d.h5:/rays of dims W x H x 8 after each."""
    from gray_amd import spacetime as S
    r = subprocess.run([J.LAUNCHER, "gray", "st=Kerr", "ic=infcam", "coor=pxcenter", "a=0.9375",
                        "n_width=12", "n_height=8", "w=48", "h=32", "r=1000", "i=60", "p=8",
                        "N=2", f"ray={tmp_path}/ray%04d.h5"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "2: -32.0 -> -64.0" in r.stdout
    assert not os.path.exists(f"{tmp_path}/ray0000.h5")
    p = O.make_params(0.9375)
    data, info = O.icond_all(p, 12, 8, 48.0, 32.0, 1000.0, 60.0, 0.0)
    for i in (1, 2):
        O.evolve_all(p, data, info, -32.0, 1024)
        rays = S.h5_read_rays(f"{tmp_path}/ray{i:04d}.h5")
        assert rays.shape == (12, 8, 8)
        assert np.array_equal(rays, data.reshape(8, 12, 8).transpose(1, 0, 2)), i
