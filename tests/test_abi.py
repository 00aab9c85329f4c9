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


def test_param_struct_layout_matches_header(tmp_path):
    """ctypes layouts equal the C compiler's for the header's structs."""
    import shutil
    import subprocess
    from neuralpom_amd import _lib
    assert len(_lib.FIELDS) == 61 and len(_lib.GEOM) == 18
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    fields = [f for f, _ in _lib.Params._fields_]
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "neuralpom_amd.h"\nint main(void){\n'
        '  printf("%zu %zu\\n", sizeof(pom_params_t), sizeof(pom_member_t));\n'
        + "".join(f'  printf("%zu\\n", offsetof(pom_params_t, {f}));\n' for f in fields)
        + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert int(out[0]) == ctypes.sizeof(_lib.Params)
    assert int(out[1]) == ctypes.sizeof(_lib.Member)
    for f, off in zip(fields, out[2:]):
        assert getattr(_lib.Params, f).offset == int(off), f


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from neuralpom_amd import _lib
    monkeypatch.setattr(_lib, "_LIB_PATH", tmp_path / "nope.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.HipExtensionError):
        _lib.load()
