"""C ABI boundary checks that need no GPU: libspdk.so loads, exports every
entry point include/spdk.h declares, and the host driver's configuration
handling (athinput parsing, overrides, error codes) behaves like the
reference's (src/main.cpp:5-189, src/parameter_input.cpp)."""
import ctypes
import os
import re

import pytest

import oracle_py as op
import spd_k_amd

HEADER = os.path.join(op.ROOT, "include", "spdk.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spdk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = spd_k_amd.lib()
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert missing == []


def test_driver_executable_built():
    assert os.access(op.DRIVER_EXE, os.X_OK)


def _run_argv(args):
    L = spd_k_amd.lib()
    argv = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
    res = spd_k_amd.RunResult()
    return L.spdk_run_argv(len(args), argv, b"/tmp/spdk_cfg_test", ctypes.byref(res))


@pytest.mark.parametrize("args", [
    ["mesh/nx1=1"],                              # x must be active (main.cpp:139-142)
    ["problem/problem=nonsense"],                # unknown problem (main.cpp:13-26)
    ["mesh/x1_bc=open"],                         # unknown boundary (main.cpp:5-11)
    ["time/integrator=rk4"],                     # unknown integrator (main.cpp:68-84)
    ["job/fallback=maybe"],                      # not a boolean (parameter_input.cpp:101-108)
    ["stray-argument"],                          # unrecognised argument (main.cpp:115-118)
    ["-i", "/nonexistent/file.athinput"],        # missing input file
])
def test_config_errors_map_to_config_code(args):
    assert _run_argv(args) == -5  # SPDK_E_CONFIG, the reference's exit(1)


def test_no_gpu_is_reported_not_faked():
    """Without a gfx950 device the context refuses to start (no CPU fallback)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    ops = {k: __import__("numpy").zeros((5, 5)) for k in spd_k_amd.Context.MATS}
    ops["wt"] = __import__("numpy").ones(4)
    with pytest.raises(RuntimeError):
        spd_k_amd.Context(3, (1, 1, 1), (0, 0, 0), 4, ops)
    assert _run_argv(["mesh/nx1=4", "mesh/nx2=4", "mesh/nx3=4", "time/tlim=0.001"]) in (-4, -5)
