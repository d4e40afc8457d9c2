"""The C-ABI library loads and exports every function include/cumetis_amd.h
declares (no GPU needed: nothing is called)."""
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "cumetis_amd.h"
LIB = ROOT / "cumetis_amd" / "libcumetis_amd.so"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*([A-Za-z_][A-Za-z_0-9]*)\s*\(", text, re.M)
    return sorted(set(names))


def test_header_parses():
    names = declared()
    assert "cm_coarsen" in names and "cuMetis_PartGraph" in names and len(names) > 20


def test_library_exports_every_declared_symbol():
    if not LIB.exists():
        pytest.fail(f"{LIB} not built")
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (\S+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_covers_header_and_loads():
    from cumetis_amd import _lib
    assert sorted(_lib.declared_symbols()) == declared()
    lib = _lib.load()
    assert lib.cm_missing_symbols == []


def test_library_has_gfx950_code():
    data = LIB.read_bytes()
    assert b"gfx950" in data
