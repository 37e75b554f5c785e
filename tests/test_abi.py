"""CPU checks of the C ABI: libjfx.so loads and exports every function that
include/jfx.h declares, and the ctypes signature table agrees with the header
(names and parameter counts). No compute call is made."""

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "jfx.h"


def _declarations():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(jfx_\w+)\s*\(([^)]*)\)\s*;", text):
        params = m.group(2).strip()
        out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


@pytest.fixture(scope="module")
def libpath():
    from jaxframes_amd import _lib, build
    if not _lib.LIB_PATH.exists():
        build.build(force=False)
    assert _lib.LIB_PATH.exists()
    return _lib.LIB_PATH


def test_header_parses():
    decls = _declarations()
    assert len(decls) >= 20
    assert decls["jfx_sort"] == 13
    assert decls["jfx_version"] == 0


def test_library_exports_every_declared_symbol(libpath):
    L = ctypes.CDLL(str(libpath))
    missing = [name for name in _declarations() if not hasattr(L, name)]
    assert not missing, missing


def test_no_undeclared_exports(libpath):
    nm = subprocess.run(["nm", "-D", "--defined-only", str(libpath)], capture_output=True, text=True)
    if nm.returncode != 0:
        pytest.skip("nm unavailable")
    exported = {ln.split()[-1] for ln in nm.stdout.splitlines() if ln.split() and ln.split()[-1].startswith("jfx_")}
    assert exported == set(_declarations()), exported ^ set(_declarations())


def test_ctypes_table_matches_header():
    from jaxframes_amd import _lib
    decls = _declarations()
    assert set(_lib.exported_symbols()) == set(decls)
    for name, args in _lib._SIGNATURES.items():
        assert len(args) == decls[name], (name, len(args), decls[name])


def test_loader_sets_signatures(libpath):
    from jaxframes_amd import _lib
    L = _lib.lib()
    assert L.jfx_version() == 1
    assert L.jfx_last_error() is not None
