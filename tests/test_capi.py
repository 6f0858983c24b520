"""CPU checks of the C ABI boundary: the library loads and exports every
entry point declared in include/thirring_hip.h (no compute calls here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "thirring_hip.h")
LIB = os.path.join(ROOT, "thirring-rhmc_amd", "libthirring_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(thr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("thr_init", "thr_dslash", "thr_dslashd", "thr_congrad", "thr_multishift",
                 "thr_qmrherm", "thr_derivs", "thr_coef", "thr_dslash_dev"):
        assert must in names


def test_library_exports_all_declared_symbols():
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_python_front_end_covers_header():
    import thirring_hip as G
    assert set(declared()) <= set(G.exported_symbols())


def test_front_end_refuses_without_library(tmp_path, monkeypatch):
    import importlib

    import thirring_hip as G
    monkeypatch.setattr(G, "LIBPATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(G, "_lib", None)
    try:
        G.lib()
    except G.ThrError:
        pass
    else:
        raise AssertionError("a missing library must raise, not fall back")
    importlib.reload(G)
