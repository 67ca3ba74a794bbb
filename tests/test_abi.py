"""The C ABI (CPU only): the library loads, exports every function include/flyhip.h declares,
and rejects bad arguments with an error code and message before touching a device."""

import ctypes
import os
import re

import pytest

from flyhip import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "flyhip.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(fz_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ("fz_create", "fz_bind", "fz_run", "fz_graph_create", "fz_graph_launch", "fz_destroy",
                 "fz_last_error", "fz_version"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared()) == set(_lib.SIGNATURES), "bindings and header disagree"
    assert lib.fz_version() == _lib.FZ_ABI_VERSION


def test_argument_errors_are_codes_not_crashes():
    lib = _lib.load()
    assert lib.fz_create(10, 0, None, None, None, 16, 0, None, None) == _lib.FZ_E_ARG
    assert b"null output" in lib.fz_last_error()
    h = ctypes.c_void_p()
    prm = _lib.FzParams(0.98, 0.005, 7.0, 0.0, 68.75, 0.275, 3.6, 22, 18)
    assert lib.fz_create(10, 0, None, None, None, 16, 0, ctypes.byref(prm), ctypes.byref(h)) == _lib.FZ_E_ARG
    assert b"indptr" in lib.fz_last_error()
    bad = _lib.FzParams(0.98, 0.005, 7.0, 0.0, 68.75, 0.275, 3.6, 200, 18)
    fake = ctypes.c_void_p(16)   # never dereferenced: the n_ref check fires first
    assert lib.fz_create(10, 0, fake, None, None, 16, 0, ctypes.byref(bad), ctypes.byref(h)) == _lib.FZ_E_ARG
    assert b"n_ref" in lib.fz_last_error()
    assert lib.fz_bind(None, None) == _lib.FZ_E_ARG
    assert lib.fz_run(None, None, 1, None, None, None, None) == _lib.FZ_E_ARG
    assert lib.fz_graph_launch(None, None) == _lib.FZ_E_ARG
    assert lib.fz_destroy(None) == _lib.FZ_OK
    with pytest.raises(ValueError):
        _lib.check(_lib.FZ_E_ARG, "probe")


def test_loaded_library_was_built_from_these_sources():
    # the library carries the hash of the sources it was compiled from; a stale build (sources
    # edited, library not rebuilt) is reported here instead of being tested silently
    info = _lib.library_info()
    assert info["build_info"].startswith("src ")
    assert info["matches_sources"], info
