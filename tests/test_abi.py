"""The C-ABI library builds for gfx950, loads, and exports every entry point
include/tmc_dmc.h declares; error paths return status codes (CPU only,
no compute call)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import ROOT


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b((?:dmc|tmc)_[a-z0-9_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    from tmc_amd import build
    if not os.path.exists(build.ENGINE_SO):
        build.build_engine()
    from tmc_amd import load_library
    return load_library()


def test_exports_every_declared_symbol(lib):
    from tmc_amd.build import ENGINE_SO
    names = _declared("tmc_dmc.h")
    assert len(names) >= 13
    out = subprocess.run(["nm", "-D", "--defined-only", ENGINE_SO], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(lib, n)   # resolvable through ctypes


def test_python_binding_lists_the_header():
    import tmc_amd
    assert sorted(tmc_amd.EXPORTS) == _declared("tmc_dmc.h")


def test_code_object_is_gfx950():
    from tmc_amd.build import ENGINE_SO
    data = open(ENGINE_SO, "rb").read()
    assert b"gfx950" in data


def test_status_strings_and_null_context(lib):
    assert lib.dmc_status_str(0) == b"ok"
    assert lib.dmc_status_str(1) == b"invalid argument"
    # null context / null outputs are rejected without touching a device
    assert lib.dmc_destroy(None) == 1
    assert lib.dmc_extract(None, None, None, 0, 0.0, 0, None, None) == 1
    assert lib.dmc_create(None, 0) == 1
    n = C.c_int(0)
    assert lib.dmc_stage_times(None, None, None, C.byref(n)) == 1
    assert lib.dmc_get_launch(None, None) == 1
    assert lib.dmc_set_launch(None, None) == 1
    assert lib.dmc_autotune(None, None, 0, 0, 0, 1, None, None, None) == 1
    assert lib.tmc_extract(None, None, None, 0, 0.0, 0, None) == 1
    assert lib.dmc_extract_noalloc(None, None, None, 0, 0.0, 0, None, None, None) == 1
    assert lib.dmc_context_bytes(None, None) == 1
    assert lib.dmc_context_trim(None) == 1


def test_no_device_is_reported(lib):
    """Without a GPU dmc_create returns DMC_ERR_NO_DEVICE (5), never exits."""
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK):
        pytest.skip("a GPU is present")
    ctx = C.c_void_p()
    assert lib.dmc_create(C.byref(ctx), 0) == 5
    assert not ctx.value


def test_cpp_wrapper_header_compiles():
    """include/tmc_dmc.hpp (the reference's C++ call surface) compiles
    against the C ABI with the host compiler."""
    src = os.path.join(ROOT, "tests", "cpp", "dropin_main.cpp")
    out = "/tmp/tmc_dropin_check.o"
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    del out
