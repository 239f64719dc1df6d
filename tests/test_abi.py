"""The C-ABI library: loads without a GPU, exports every entry point of include/cltem_amd.h,
and the host-only functions behave (CPU only; no compute call needs a GPU here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import cltem_amd
from cltem_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cltem_amd.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mse_\w+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    lib = _abi.load()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (mse_\w+)", out))
    assert set(names) <= exported


def test_python_binding_covers_header():
    assert set(header_functions()) == set(_abi.EXPORTED)


def test_library_has_gfx950_code_object():
    # the fat binary embedded by hipcc names its offload targets
    blob = open(_abi.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_plan_geometry_rejects_bad_input():
    s = cltem_amd.make_scenario(resolution=0, struct_min=(0, 0, 0), struct_max=(1, 1, 1),
                                area=(0, 1, 0, 1))
    with pytest.raises(cltem_amd.MseError):
        cltem_amd.plan_geometry(s)
    lib = _abi.load()
    assert lib.mse_plan_geometry(None, None) == -1


def test_engine_calls_fail_loudly_on_null_handle():
    lib = _abi.load()
    assert lib.mse_synchronize(None) != 0
    assert b"NULL" in lib.mse_last_error() or b"engine" in lib.mse_last_error()
    assert lib.mse_set_checkpoint_budget(None, 0) != 0
    assert b"NULL" in lib.mse_last_error()


def test_engine_creation_without_gpu_is_an_error():
    try:
        e = cltem_amd.Engine(0)
    except cltem_amd.MseError as err:
        assert err.code in (-1, -3)
        return
    e.close()
    pytest.skip("a GPU is present")
