"""The C-ABI library loads and exports every entry point include/*.h declares
(CPU-only: no compute call is made)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "neuralpom_amd" / "libneuralpom_amd.so"


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        names |= set(re.findall(r"^int\s+(pom_\w+)\s*\(", h.read_text(), re.M))
    return names


def test_header_declares_entry_points():
    assert {"pom_batch_create", "pom_explicit_step", "pom_explicit_step_adj"} <= _declared()


@pytest.mark.skipif(not LIB.exists(), reason="library not built (python -m neuralpom_amd.build)")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_every_declared_symbol():
    from neuralpom_amd import _lib
    assert set(_lib.EXPORTED) == _declared()


def test_param_struct_layout_matches_header():
    from neuralpom_amd import _lib
    # 22 doubles, 6 int32, 1 double (8-byte aligned) = 208 bytes
    assert ctypes.sizeof(_lib.Params) == 22 * 8 + 6 * 4 + 8
    assert ctypes.sizeof(_lib.Member) == 24
    assert len(_lib.FIELDS) == 61 and len(_lib.GEOM) == 18


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from neuralpom_amd import _lib
    monkeypatch.setattr(_lib, "_LIB_PATH", tmp_path / "nope.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.HipExtensionError):
        _lib.load()
