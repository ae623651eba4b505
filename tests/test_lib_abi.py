"""The C ABI library loads and exports every symbol the public header declares.

CPU-only: loading the library does not touch the GPU, and no compute entry point
is called here.
"""

from __future__ import annotations

import ctypes

import pytest

from numbagg_amd import _build, _lib


@pytest.fixture(scope="module")
def lib():
    if _build.library_is_stale():
        _build.build_library()
    return _lib.load()


def test_header_symbols_are_exported(lib):
    declared = _lib.header_symbols()
    assert len(declared) >= 7
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, f"declared in include/numbagg_amd.h but not exported: {missing}"


def test_bindings_cover_header():
    # Every declared entry point has a ctypes signature, and vice versa.
    assert set(_lib.header_symbols()) == set(_lib._SIGNATURES)


def test_version_and_error_strings(lib):
    assert b"gfx950" in lib.nbg_version()
    assert isinstance(lib.nbg_last_error(), bytes)


def test_workspace_queries_without_gpu(lib):
    assert lib.nbg_fill_workspace_bytes(4, 10, _lib.NBG_F32) == 0  # single tile: no look-back
    assert lib.nbg_fill_workspace_bytes(4, 100_000, _lib.NBG_F32) > 0


def test_invalid_arguments_rejected_before_launch(lib):
    # Argument checks run before any device work, so they are CPU-testable.
    rc = lib.nbg_ffill(_lib.NBG_F32, None, None, 1, 10, 10, 10, -1, None, 0, None)
    assert rc == 1
    assert b"limit" in lib.nbg_last_error()
    rc = lib.nbg_ffill(7, ctypes.c_void_p(16), ctypes.c_void_p(16), 1, 10, 10, 10, 1, None, 0, None)
    assert rc == 1
    assert b"dtype" in lib.nbg_last_error()


def test_injected_error(lib, monkeypatch):
    monkeypatch.setenv("NBG_INJECT_ERROR", "nbg_ffill")
    rc = lib.nbg_ffill(_lib.NBG_F32, None, None, 1, 10, 10, 10, 1, None, 0, None)
    assert rc == 3
    assert b"injected" in lib.nbg_last_error()
