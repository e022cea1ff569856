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
    assert L.fsg_abi_version() == N.ABI_VERSION == 20
    assert L.fsg_launch_count() >= 0


def test_struct_sizes_match_header_layout():
    # fsg_output: ptr + i64 + 6*i32-sized fields + 2 doubles + ptr
    assert ctypes.sizeof(N.fsg_window) == 24
    assert ctypes.sizeof(N.fsg_region) == 16
    assert ctypes.sizeof(N.fsg_output) == 72
    assert ctypes.sizeof(N.fsg_ray_sample) == 24
    assert ctypes.sizeof(N.fsg_ao_sample) == 32
    assert ctypes.sizeof(N.fsg_mode_stats) == 48


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


def test_ctypes_layout_matches_the_c_header(tmp_path):
    """gcc compiles include/fsg.h and prints sizeof/offsetof of every field of
    every struct the ctypes mirror declares; both must agree."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = {n: getattr(N, n) for n in dir(N)
               if n.startswith("fsg_") and isinstance(getattr(N, n), type)
               and issubclass(getattr(N, n), ctypes.Structure)}
    assert {"fsg_window", "fsg_output", "fsg_topousm_params"} <= set(structs)
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fsg.h"', "int main(void){"]
    for name, st in structs.items():
        lines.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        for f in st._fields_:
            lines.append(f'printf("{name}.{f[0]} %zu\\n", offsetof({name}, {f[0]}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    inc = os.path.join(os.path.dirname(N.__file__), "..", "include")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in
               subprocess.run([str(exe)], check=True, capture_output=True, text=True)
               .stdout.splitlines())
    for name, st in structs.items():
        assert int(got[f"{name} size"]) == ctypes.sizeof(st), name
        for f in st._fields_:
            assert int(got[f"{name}.{f[0]}"]) == getattr(st, f[0]).offset, f"{name}.{f[0]}"
