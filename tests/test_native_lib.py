"""The C-ABI library loads and exports every function declared in include/nfx.h
(no compute call: runs without a GPU)."""

import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "nfx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nfx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_table():
    from nfempy_amd import _native

    assert set(declared_functions()) == set(_native.SIGNATURES)


def test_library_exports_every_declared_symbol():
    from nfempy_amd import _native

    if not _native.library_available():
        pytest.fail("libnfx.so is not built (python -m nfempy_amd.build_lib)")
    lib = _native.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.nfx_version() >= _native.MIN_VERSION == 500


def test_errors_are_reported_without_a_gpu():
    """Argument errors come back as status codes with a message, not exceptions."""
    from nfempy_amd import _native

    lib = _native.lib()
    rc = lib.nfx_stage_reduce(-1, None, None, 0, None, None, None)
    assert rc == _native.NFX_E_INVALID
    assert b"n_seg" in lib.nfx_last_error()
    with pytest.raises(ValueError):
        _native.check(rc, "nfx_stage_reduce")
