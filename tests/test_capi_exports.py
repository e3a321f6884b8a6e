"""The C-ABI library loads on a host without a GPU and exports every function
include/cs_stitch.h declares (no compute calls)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "cs_stitch.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(cs_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from cs_stitch import _ffi
    lib = _ffi.load()
    names = declared_functions()
    assert len(names) >= 17
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ffi_signatures_cover_header():
    from cs_stitch import _ffi
    assert set(declared_functions()) <= set(_ffi.SIGNATURES)


def test_host_only_calls_work_without_gpu():
    from cs_stitch import _ffi
    lib = _ffi.load()
    assert lib.cs_version().startswith(b"crossstage")
    si = _ffi.ScaleInfo()
    assert lib.cs_make_scale_info(1920, 1080, 0.5, ctypes.byref(si)) == 0
    assert (si.work_w, si.work_h) == (960, 540)
    assert lib.cs_make_scale_info(10, 10, 1.5, ctypes.byref(si)) == 1
    n = ctypes.c_size_t()
    assert lib.cs_s2_workspace_bytes(1920, 1080, ctypes.byref(_ffi.S2Config(20, 5000, 0.04)), ctypes.byref(n)) == 0
    assert n.value > 1920 * 1080
    assert lib.cs_s2_workspace_bytes(1920, 1080, ctypes.byref(_ffi.S2Config(0, 5000, 0.04)), ctypes.byref(n)) == 1
    assert lib.cs_s4_workspace_bytes(10, 10, ctypes.byref(_ffi.S4Config(1.0, 0)), ctypes.byref(n)) == 1
