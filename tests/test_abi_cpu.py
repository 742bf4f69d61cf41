"""The C-ABI libraries load without a GPU and export every function the
public headers (include/*.h) declare; the Python bindings name the same set.
No compute call is made."""
import ctypes as C
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gray_amd")

DECL = re.compile(r"^\s*(?:const\s+)?[a-z_0-9]+(?:\s+|\s*\*+\s*)\**((?:grayhip|gray)_\w+)\s*\(",
                  re.M)


def declared():
    names = {}
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)       # drop comments
        for m in DECL.finditer(text):
            names[m.group(1)] = os.path.basename(h)
    return names


def library_for(name):
    if name.startswith("grayhip_"):
        return "libgrayhip.so"
    if name.startswith("gray_h5_"):
        return "libgray_h5.so"
    return "libgray.so"


def test_headers_declare_the_expected_surface():
    names = declared()
    for must in ("grayhip_create", "grayhip_render", "grayhip_spacetime_upload_rho",
                 "grayhip_read_bl", "grayhip_server_submit", "gray_mkmod", "gray_run",
                 "gray_h5_read_level", "gray_h5_write_rays", "gray_serve_run", "gray_bks_sample"):
        assert must in names, must
    assert len(names) > 60


@pytest.mark.parametrize("lib", ["libgrayhip.so", "libgray_h5.so", "libgray.so"])
def test_library_exports_every_declared_symbol(lib):
    path = os.path.join(PKG, lib)
    assert os.path.exists(path), f"{path} not built"
    L = C.CDLL(path)
    missing = [n for n in declared() if library_for(n) == lib and not hasattr(L, n)]
    assert not missing, missing


def test_python_bindings_cover_the_headers():
    from gray_amd import grayhip, job, serve, spacetime
    bound = set(grayhip.EXPORTS) | set(spacetime.EXPORTS) | set(serve.EXPORTS) | set(job.EXPORTS)
    names = set(declared())
    assert names - bound == set(), sorted(names - bound)
    assert bound - names == set(), sorted(bound - names)
