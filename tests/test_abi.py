"""CPU: the C-ABI library loads and exports every entry point include/*.h declares,
the Python binding declares the same set, and calls fail loudly (no CPU
fallback) when no device is present."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(bh_[a-z0-9_]+)\s*\(", txt))
    return names


def test_header_declares_entry_points():
    names = _declared()
    for must in ("bh_dgbmv", "bh_sgbmv", "bh_dgbmv_multi", "bh_dgbmm", "bh_dgbtrf", "bh_dgbtrs", "bh_dgbpad"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from banded_hip import _lib

    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(_declared()) if not hasattr(L, n)]
    assert not missing, missing


def test_binding_signatures_cover_header():
    from banded_hip import _lib

    assert set(_lib.SIGNATURES) == _declared()


def test_no_device_fails_loudly():
    """Without a GPU every call reports an error instead of computing on the CPU."""
    from banded_hip import _lib

    L = _lib.load()
    cnt = ctypes.c_int(-1)
    assert L.bh_device_count(ctypes.byref(cnt)) == 0
    if cnt.value > 0:
        pytest.skip("a device is visible")
    assert L.bh_init(0) == -1002
    p = ctypes.c_void_p()
    assert L.bh_malloc(ctypes.byref(p), 64) != 0
    assert b"device" in L.bh_last_error()


def test_argument_errors_are_lapack_style():
    """Invalid arguments return -i (1-based position) before touching the device."""
    from banded_hip import _lib

    L = _lib.load()
    assert L.bh_dgbmv(b"X", 1, 1, 0, 0, 1.0, None, 1, None, 1, 0.0, None, 1) == -1
    assert L.bh_dgbmv(b"N", -1, 1, 0, 0, 1.0, None, 1, None, 1, 0.0, None, 1) == -2
    assert L.bh_dgbmv(b"N", 1, 1, 2, 2, 1.0, None, 3, None, 1, 0.0, None, 1) == -8
    info = ctypes.c_int64()
    assert L.bh_dgbtrf(4, 4, 1, 1, None, 3, None, ctypes.byref(info)) == -6
    assert L.bh_dgbtrs(b"N", 4, 1, 1, 1, None, 4, None, None, 2) == -10


def test_package_import_without_library_is_loud(tmp_path, monkeypatch):
    from banded_hip import _lib

    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError):
        _lib.load()
