"""The C-ABI library loads and exports every symbol include/fsg.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os

import pytest

from fujishadergpu_amd import _native as N


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(N.LIB_PATH)
    declared = N.declared_symbols()
    assert len(declared) >= 15
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, f"libfsg.so lacks {missing}"


def test_binding_table_covers_header():
    # every declared entry point has a ctypes signature
    assert set(N.declared_symbols()) <= set(N._SIGS)


def test_loader_validates_abi_and_symbols():
    L = N.lib()
    assert L.fsg_abi_version() == N.ABI_VERSION == 2
    assert L.fsg_launch_count() >= 0


def test_struct_sizes_match_header_layout():
    # fsg_output: ptr + i64 + 6*i32-sized fields + 2 doubles + ptr
    assert ctypes.sizeof(N.fsg_window) == 24
    assert ctypes.sizeof(N.fsg_region) == 16
    assert ctypes.sizeof(N.fsg_output) == 72
    assert ctypes.sizeof(N.fsg_ray_sample) == 24
    assert ctypes.sizeof(N.fsg_ao_sample) == 32


def test_invalid_arguments_fail_without_gpu():
    L = N.lib()
    w = N.fsg_window(0, 0, -1, 5)
    r = N.fsg_region(0, 0, 1, 1)
    rc = L.fsg_copy_region(ctypes.byref(w), ctypes.byref(r), None, None)
    assert rc == N.FSG_E_INVALID
    assert b"bad window" in L.fsg_last_error()
    with pytest.raises(N.FsgError):
        N.check(rc)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    import importlib
    monkeypatch.setenv("FSG_LIBRARY", str(tmp_path / "nope.so"))
    mod = importlib.reload(N)
    try:
        with pytest.raises(mod.NativeLibraryError):
            mod.lib()
    finally:
        monkeypatch.delenv("FSG_LIBRARY")
        importlib.reload(N)
