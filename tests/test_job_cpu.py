"""Job surface on the CPU: option tables, conf return codes, raw dump
files, and clean failure without a GPU (sim-org/gray.c:62-111, io.c:39-72)."""
import math
import os

import numpy as np
import pytest

from gray_amd import job as J

# (option, alias, default) rows of sim-org/icond.opts, param.opts,
# setup.opts and sim/Kerr.opts
DEFAULTS = [
    ("w_img", "w", 64.0), ("h_img", "h", 64.0), ("r_obs", "r", 1024.0),
    ("i_obs", "i", 60.0), ("j_obs", "j", 0.0), ("M_ADM", "m", 4.3e6),
    ("coordinates", "coor", "dyst"), ("dyst_file", "data", "data.h5"),
    ("enable_fast_light", "fast_light", 0), ("w_rays", "W", 512), ("h_rays", "H", 512),
    ("n_electron", "ne", 1e6), ("beta_crit", "bc", 1.0), ("R_high", "Rh", 1.0),
    ("R_low", "Rl", 1.0), ("i_platform", "ip", 0), ("i_device", "id", 0),
    ("precision", "p", 4), ("scheme", "scm", "RK4"), ("morder", "mo", "AoS"),
    ("kflags", "kf", "-w"), ("outfile", "out", "%04d.raw"), ("t_init", "t0", 0.0),
    ("dt_dump", "dt", -32.0), ("i_init", "i0", 0), ("n_sub", "n", 1024),
    ("n_dump", "N", 64), ("m_bh", "M", 4e6), ("a_spin", "a", 0.0),
    ("b_field", "bz", 0.0),    # build addition: the thermal flow's field b^z
]


def val(s, ref):
    if isinstance(ref, str):
        return s
    return float(s)


@pytest.mark.parametrize("name,alias,default", DEFAULTS)
def test_defaults_and_aliases(name, alias, default):
    with J.Job() as j:
        assert val(j.get(name), default) == default
        assert j.get(alias) == j.get(name)
        num = 8 if name == "precision" else 7
        new = str(num) if not isinstance(default, str) else "XYZ"
        assert j.conf(f"{alias}={new}") == 0
        if isinstance(default, str):
            assert j.get(name) == "XYZ"
        else:
            assert float(j.get(name)) == num


def test_conf_return_codes():
    with J.Job() as j:
        assert j.conf("nosuch=1") != 0
        assert j.conf("w_img") != 0            # no '='
        assert j.conf("w_img=abc") != 0        # not a number
        assert j.conf("W=-3") != 0             # size_t
        assert j.conf("p=3") != 0              # precision is 4 or 8
        assert j.conf("p=double") == 0 and j.get("p") == "8"
        assert j.conf("precision=single") == 0 and j.get("p") == "4"
        assert j.conf("mo=SoA") == 0 and j.get("morder") == "SoA"
        assert j.conf("ml=AoS") == 0 and j.get("morder") == "AoS"
        assert j.conf("type=GPU") == 0 and j.conf("type=CL_DEVICE_TYPE_CPU") == 0
        assert j.conf("type=TPU") != 0


def test_nu_list_sets_channels():
    with J.Job() as j:
        assert j.n_freq == 0 and j.n_data == 8
        assert j.conf("nu=1,2.67,4,8") == 0
        assert j.n_freq == 4 and j.n_data == 16
        assert [float(x) for x in j.get("nu").split(",")] == [1.0, 2.67, 4.0, 8.0]
        assert j.conf("nu=[5e11, 2.3e11]") == 0 and j.n_freq == 2
        assert j.conf("nu=") == 0 and j.n_freq == 0
        assert j.conf("nu=1,x") != 0


def test_rewrite_names_map_to_the_same_options():
    with J.Job() as j:
        assert j.conf("n_width=33") == 0 and j.get("w_rays") == "33"
        assert j.conf("n_height=17") == 0 and j.get("H") == "17"
        assert j.conf("st=Kerr") == 0 and j.get("spacetime") == "Kerr"


def test_raw_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    for t in (np.float64, np.float32):
        w, h, n = 5, 3, 16
        data = rng.normal(size=(h * w, n)).astype(t)
        info = rng.normal(size=h * w).astype(t)
        p = str(tmp_path / f"x{np.dtype(t).itemsize}.raw")
        J.write_raw(p, data, info, w, h)
        s, d = J.load_raw(p)
        assert s.dtype == t and s.shape == (h, w, n) and d.shape == (h, w)
        assert np.array_equal(s.reshape(h * w, n), data) and np.array_equal(d.ravel(), info)
        assert os.path.getsize(p) == 4 * 8 + data.nbytes + info.nbytes


def test_launcher_rejects_bad_options():
    import subprocess
    r = subprocess.run([J.LAUNCHER, "gray", "bogus=1"], capture_output=True, text=True)
    assert r.returncode != 0 and "unknown option" in r.stderr
    r = subprocess.run([J.LAUNCHER, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "a_spin:a" in r.stdout


def test_init_fails_cleanly_on_bad_setup():
    with J.Job() as j:
        assert j.conf("scm=RK45") == 0
        assert j.init() != 0
    with J.Job() as j:
        assert j.conf("mo=XoY") == 0
        assert j.init() != 0
    with J.Job() as j:
        assert j.exec() != 0  # exec before init


def test_ray_file_layout_and_exclusive_create(tmp_path):
    """sim/gray.c:117,176: /rays holds palloc(real, W, H, 8), last index
    fastest; the file is created with H5F_ACC_EXCL."""
    from gray_amd import spacetime as S
    rng = np.random.default_rng(4)
    states = rng.normal(size=(3, 5, 8))          # [H, W, 8] dump order
    p = str(tmp_path / "ray0001.h5")
    S.h5_write_rays(p, states)
    got = S.h5_read_rays(p)
    assert got.shape == (5, 3, 8)
    assert np.array_equal(got, states.transpose(1, 0, 2))
    with pytest.raises(S.H5Error):
        S.h5_write_rays(p, states)               # exists
    p4 = str(tmp_path / "ray0002.h5")
    S.h5_write_rays(p4, states, dtype=4)
    assert np.array_equal(S.h5_read_rays(p4), states.transpose(1, 0, 2).astype(np.float32))


@pytest.mark.parametrize("args,msg", [
    (["st=Schwarzschild"], "Unknown spacetime configuration"),
    (["st=Kerr", "ic=plane"], "Unknown initial conditions"),
    (["st=Kerr", "nu=1e11"], "8 reals per ray"),
    (["st=Kerr", "coor=polar"], "pxcenter"),
])
def test_rewrite_surface_rejects(args, msg):
    import subprocess
    r = subprocess.run([J.LAUNCHER, "gray", "W=4", "H=4"] + args, capture_output=True, text=True)
    assert r.returncode != 0 and msg in r.stderr, r.stderr
