"""The C-ABI library loads and exports every symbol include/saige_hip.h
declares, and the ctypes table covers the same set (no compute calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "saige_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(saige_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    from saige_gpu import _lib
    assert os.path.exists(_lib.LIB_PATH), "build libsaigehip.so first (make -C saige-gpu_amd)"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_load_sets_signatures():
    from saige_gpu import _lib
    L = _lib.load()
    assert L.saige_last_error.restype is ctypes.c_char_p
    # a null context is rejected without touching a device
    assert L.saige_get_nnomissing(None) == 0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from saige_gpu import _lib
    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    try:
        _lib.load()
    except RuntimeError as e:
        assert "not built" in str(e)
    else:
        raise AssertionError("load() must raise when the library is missing")


def test_serving_entry_points_reject_null_handles():
    """saige_server_start / saige_submit / saige_poll / saige_server_stop
    refuse null handles with SAIGE_E_ARG before touching a device."""
    from saige_gpu import _lib
    L = _lib.load()
    h = ctypes.c_void_p()
    assert L.saige_server_start(None, None, 4, 0, ctypes.byref(h)) == _lib.SAIGE_E_ARG
    spec = _lib.FitSpec()
    t = ctypes.c_int64(0)
    assert L.saige_submit(None, ctypes.byref(spec), ctypes.byref(t)) == _lib.SAIGE_E_ARG
    res = _lib.FitResult()
    assert L.saige_poll(None, 1, 0, ctypes.byref(res), None) == _lib.SAIGE_E_ARG
    assert L.saige_server_stop(None) == _lib.SAIGE_E_ARG
