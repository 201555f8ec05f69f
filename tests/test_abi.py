"""The C-ABI library loads and exports every entry point include/scopi_hip.h declares.

No compute call is made here (no GPU in this test)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    with open(os.path.join(ROOT, "include", "scopi_hip.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(scopi_hip_\w+)\s*\(", text, re.M)))


def test_header_lists_entry_points():
    names = declared()
    assert "scopi_hip_run" in names and "scopi_hip_solve" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    from scopi_amd import hip
    path = hip.LIB_PATH
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(hip.EXPORTS) == declared()


def test_python_binding_refuses_missing_library(monkeypatch, tmp_path):
    from scopi_amd import hip
    monkeypatch.setattr(hip, "_lib", None)
    monkeypatch.setattr(hip, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(hip.ScopiHipError):
        hip.lib()
