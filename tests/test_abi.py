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


def _param_table(path):
    """block/name -> value of an input file (comments and blank lines skipped)."""
    table, block = {}, None
    for raw in open(path):
        s = raw.split("#", 1)[0].strip()
        if not s:
            continue
        if s.startswith("<"):
            block = s[1:s.index(">")].strip()
        else:
            k, v = s.split("=", 1)
            table[f"{block}/{k.strip()}"] = v.strip()
    return table


def test_parameter_dump_is_a_loadable_input_with_last_assignment_winning(tmp_path):
    """The effective parameters written with outputs on (src/main.cpp:248-253) reload to the same
    values; a later file line beats an earlier one and a command-line assignment beats the file
    (src/parameter_input.cpp:24-56)."""
    src = tmp_path / "in.athinput"
    src.write_text("<mesh>\nnx1 = 6   # first\nnx2 = 4\n<time>\ntlim = 0.5\n<mesh>\nnx1 = 7\n"
                   "<output>\ndt = 0.1\n")
    out1, out2 = tmp_path / "o1", tmp_path / "o2"
    args = ["-i", str(src), "mesh/nx2=5", "time/cfl = 0.25", "time/nlim=1"]
    for out, a in ((out1, args), (out2, None)):
        if a is None:
            a = ["-i", str(out1 / "parameters.txt")]
        L = spd_k_amd.lib()
        argv = (ctypes.c_char_p * len(a))(*[x.encode() for x in a])
        rc = L.spdk_run_argv(len(a), argv, str(out).encode(), None)
        assert rc in (0, -4, -5), rc  # without a GPU the run stops after the dump, before any step
    t1 = _param_table(out1 / "parameters.txt")
    assert t1["mesh/nx1"] == "7" and t1["mesh/nx2"] == "5" and t1["time/tlim"] == "0.5"
    assert float(t1["time/cfl"]) == 0.25 and float(t1["hydro/gamma"]) == 1.4  # a default is recorded
    assert _param_table(out2 / "parameters.txt") == t1


@pytest.mark.parametrize("text", [
    "<mesh\nnx1 = 4\n",          # header without '>'
    "<  >\nnx1 = 4\n",           # header naming no block
    "nx1 = 4\n",                 # assignment before any block
    "<mesh>\nnx1\n",             # neither header nor assignment
    "<mesh>\nnx1 =   # none\n",  # empty value
    "<mesh>\n = 4\n",            # empty name
    "<mesh>\nnx1 = four\n",      # not an integer where one is read
])
def test_input_file_syntax_errors_are_config_errors(tmp_path, text):
    f = tmp_path / "bad.athinput"
    f.write_text(text)
    assert _run_argv(["-i", str(f)]) == -5
