"""CPU checks of the product library: it loads without a GPU and exports
every function include/libdwt.h declares (no compute calls here)."""
import ctypes
import os
import subprocess

import pytest

import libdwt_amd


def test_library_built():
    assert os.path.exists(libdwt_amd.LIB_PATH), "build with python -m libdwt_amd.build"


def test_exports_every_declared_symbol():
    names = libdwt_amd.header_functions()
    assert len(names) >= 45
    lib = libdwt_amd.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_header_covers_reference_drop_in_set():
    """The reference's 1D, zero-padding and 2D entry points (src/libdwt.h:22-592)."""
    names = set(libdwt_amd.header_functions())
    for w in ("cdf97", "cdf53"):
        for p in ("s", "d"):
            for f in (f"dwt_{w}_2f_{p}", f"dwt_{w}_2i_{p}", f"dwt_{w}_f_ex_stride_{p}",
                      f"dwt_{w}_i_ex_stride_{p}", f"dwt_{w}_f_ex_{p}", f"dwt_{w}_i_ex_{p}",
                      f"dwt_{w}_f_{p}", f"dwt_{w}_i_{p}"):
                assert f in names, f
    for p in ("s", "d"):
        for f in ("f", "i"):
            assert f"dwt_zero_padding_{f}_{p}" in names
            assert f"dwt_zero_padding_{f}_stride_{p}" in names


def test_exported_symbols_are_c_abi():
    out = subprocess.run(["nm", "-D", "--defined-only", libdwt_amd.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for n in libdwt_amd.header_functions():
        assert n in exported, f"{n} not exported unmangled"


def test_header_compiles_as_c():
    src = '#include "libdwt.h"\nint main(void){return 0;}\n'
    inc = os.path.dirname(libdwt_amd.HEADER)
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", inc, "-x", "c", "-",
                        "-o", os.devnull], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_extension_reports_bad_arguments_without_gpu_compute():
    lib = libdwt_amd.lib()
    j = ctypes.c_int(-1)
    rc = lib.dwt_gpu_2d_async(7, 0, 0, None, 4, 4, 4, 4, 4, 4, ctypes.byref(j), 0, 0, None)
    assert rc == -1
    assert b"wavelet" in lib.dwt_gpu_last_error()
