"""The C-ABI library loads and exports every function include/dtc_capi.h declares
(no compute call here: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "dtc_capi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dtc_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    import dtc_amd._lib as L
    lib = ctypes.CDLL(L.LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_declares_every_symbol():
    import dtc_amd._lib as L
    assert set(_declared()) == set(L.SIGNATURES)


def test_struct_layouts_match_the_header():
    import dtc_amd._lib as L
    # sizes implied by the header (all members naturally aligned, no implicit padding)
    assert ctypes.sizeof(L.DomainDesc) == 8 + 4 * 4 + 3 * 8 + 8 + 3 * 8 + 3 * 8 + 8 + 3 * 8 + 8
    assert ctypes.sizeof(L.SolveOpts) == 4 + 4 + 8 + 8 + 8 + 8 + 8 + 4 + 4
    assert ctypes.sizeof(L.DomainResult) == 8 + 6 * 8 + 3600 * 16 + 72 * 8 + 8 + 2 * 8 + 6 * 8 + 18 * 8 + 72 * 8 + 2 * 8 + 2 * 8


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import importlib

    import dtc_amd._lib as L
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(L, "_lib", None)
    try:
        L.load()
    except ImportError as e:
        assert "missing" in str(e)
    else:
        raise AssertionError("load() must raise without the HIP library")
    importlib.reload(L)
