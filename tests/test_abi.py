"""C-ABI library: it loads without a GPU and exports every symbol the header
declares (no compute calls here)."""
import re
import subprocess
from pathlib import Path

import pytest

from projectracer_amd import _abi

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "projectracer_gpu.h"


def _declared():
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pt_[a-z_]+)\s*\(", txt)))


def test_header_declares_the_python_binding_set():
    assert _declared() == sorted(_abi.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    if not _abi.LIB_PATH.exists():
        pytest.skip("libprojectracer_gpu.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", str(_abi.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (pt_[a-z_]+)$", out, flags=re.M))
    missing = [s for s in _declared() if s not in exported]
    assert not missing, missing


def test_library_loads_and_reports_no_device_cleanly():
    if not _abi.LIB_PATH.exists():
        pytest.skip("libprojectracer_gpu.so not built")
    lib = _abi.load_library()
    p = _abi.PtParams()
    lib.pt_default_params(p)
    assert p.nstageS == 35 and p.grids == 6 and abs(p.tolerance - 1e-8) < 1e-20
    # invalid arguments are rejected before any device work
    bad = _abi.default_params(nx=2)
    import ctypes as C
    ctx = C.c_void_p()
    assert lib.pt_create(C.byref(bad), C.byref(ctx)) == _abi.PT_ERR_INVALID
    assert b"nx" in lib.pt_last_error()
