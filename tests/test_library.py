"""The C-ABI library loads and exports every entry point include/flusi_hip.h
declares (no GPU call is made).  CPU only."""
import ctypes
import os
import re

import pytest

import flusi_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "flusi_hip.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(flusi_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_boundary():
    names = _declared()
    for must in ("flusi_init", "flusi_cal_nlk", "flusi_fluid_timestep", "flusi_create_mask",
                 "flusi_write_integrals", "flusi_cal_vis", "flusi_fft", "flusi_ifft"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(flusi_amd.LIB_PATH):
        pytest.fail("libflusi_hip.so is not built (python -m flusi_amd.build)")
    lib = ctypes.CDLL(flusi_amd.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares exactly the header's entry points
    assert sorted(flusi_amd.EXPORTED) == _declared()


def test_binding_fails_loudly_without_library(tmp_path):
    with pytest.raises(RuntimeError, match="missing"):
        flusi_amd.load_library(str(tmp_path / "nope.so"))


def test_fortran_interface_module_links():
    """The iso_c_binding module and the Fortran driver build against the
    library (amdflang from ROCm); no GPU call is made."""
    from flusi_amd import build as b
    if not os.path.exists(b.FLANG):
        pytest.skip("no Fortran compiler in this image")
    if not os.path.exists(flusi_amd.LIB_PATH):
        pytest.fail("libflusi_hip.so is not built")
    os.makedirs(b.BUILD, exist_ok=True)
    exe = b.build_fortran()
    assert os.path.exists(exe)
    import subprocess
    r = subprocess.run([exe], capture_output=True, text=True)
    assert "usage: flusi_amd" in r.stdout
