"""The C ABI library loads and exports every entry point declared in
include/lmhip.h; calls that need no device fail cleanly (CPU only)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lmhip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lmh_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from lmhip import _lib
    lib = ctypes.CDLL(_lib.LIBPATH)
    names = declared()
    assert "lmh_step" in names and "lmh_update_solver" in names and len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from lmhip import _lib
    L = _lib.lib()
    for n in declared():
        assert getattr(L, n).restype is not None or n == "lmh_last_error"


def test_null_arguments_are_rejected_without_device():
    from lmhip import _lib
    L = _lib.lib()
    assert L.lmh_version() >= 100
    assert L.lmh_set_geometry(None, 4, 0, None, None, None, None, None, None) == _lib.LMH_EINVAL
    assert b"null context" in L.lmh_last_error()
    assert L.lmh_step(None, None) == _lib.LMH_EINVAL
    assert L.lmh_state_upload(None, None, 1) == _lib.LMH_EINVAL
    assert L.lmh_get_stream(None, None) == _lib.LMH_EINVAL
    assert L.lmh_set_progress(None, _lib.PROGRESS_FN(), None) == _lib.LMH_EINVAL


def test_no_cpu_fallback_without_device():
    """Creating a context on a machine without a HIP device raises; nothing
    silently computes on the CPU."""
    from lmhip import _lib
    n = ctypes.c_int(-1)
    rc = _lib.lib().lmh_device_count(ctypes.byref(n))
    if rc == 0 and n.value > 0:
        pytest.skip("a HIP device is present")
    import lmhip
    with pytest.raises((lmhip.LMHError, ValueError)):
        lmhip.Context(0)
