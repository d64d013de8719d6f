"""The C-ABI library loads here (no GPU) and exports every symbol the header declares."""

import os
import re

import pytest

from gradwave_amd import _lib as L

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "gradwave_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gw_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "gw_hpsi_apply" in names and "gw_zgemm_inner" in names and len(names) >= 15


@pytest.mark.skipif(not os.path.exists(L.LIB_PATH), reason="library not built")
def test_library_exports_all_declared_symbols():
    lib = L.lib()
    for name in declared_functions():
        assert hasattr(lib, name), f"{name} declared in gradwave_hip.h but not exported"
        assert name in L.SIGNATURES, f"{name} has no ctypes signature in _lib.py"
    assert lib.gw_version() >= 1


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(L.GWError, match="not built"):
        L.lib()
