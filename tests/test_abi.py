"""The C-ABI libraries load and export every symbol their headers declare
(no compute calls: runs without a GPU)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(shci_[a-z0-9_]+)\s*\(", text)))


def test_hip_lib_exports_header_symbols():
    from shci_amd import _abi, _native

    lib = _native.hip_lib()
    names = _declared("shci_hip.h")
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
        assert name in _abi.EXPORTS, f"{name} missing from the ctypes declarations"


def test_host_lib_exports_header_symbols():
    from shci_amd import _native

    lib = _native.host_lib()
    for name in _declared("shci_host.h"):
        assert hasattr(lib, name), name


def test_no_cpu_fallback_when_library_missing(tmp_path, monkeypatch):
    from shci_amd import _native

    monkeypatch.setattr(_native, "_HIP", tmp_path / "missing.so")
    monkeypatch.setattr(_native, "_hip", None)
    try:
        _native.hip_lib()
    except RuntimeError as e:
        assert "missing" in str(e)
    else:
        raise AssertionError("expected a hard failure")
