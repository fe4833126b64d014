"""CPU checks of the C-ABI boundary: the library loads and exports exactly
what include/jfx.h declares, with the signatures the binding uses."""

import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "jfx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(jfx_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from jaxfun_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libjfx.so not built")
    lib = _lib.load()
    decl = _declared()
    assert decl, "no declarations parsed"
    for name in decl:
        assert hasattr(lib, name), name
    assert sorted(_lib.SIGNATURES) == decl
    assert lib.jfx_version() >= 100


def test_header_argument_counts_match_binding():
    from jaxfun_amd import _lib

    src = open(os.path.join(ROOT, "include", "jfx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    for name, (_, args) in _lib.SIGNATURES.items():
        m = re.search(name + r"\s*\(([^)]*)\)", src)
        assert m, name
        params = [p for p in m.group(1).split(",") if p.strip() and p.strip() != "void"]
        assert len(params) == len(args), (name, params, args)
