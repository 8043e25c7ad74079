"""Builds libspdk.so (gfx950 HIP kernels + C ABI + host driver) and the
`spdk` command-line driver into spd_k_amd/_build/.

HIP sources are compiled with hipcc for gfx950 only; host C++ with g++.
Both use -ffp-contract=off so every multiply and add rounds separately,
which the FP64 path needs to reproduce the reference arithmetic bit for bit.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_SOURCES = ["csrc/sd_stages.hip", "csrc/fv_stages.hip", "csrc/fused.hip", "csrc/induction_stages.hip",
               "csrc/induction_fused.hip", "csrc/mixed.hip", "csrc/mlp.hip"]
HOST_SOURCES = ["csrc/host/params.cpp", "csrc/host/setup.cpp", "csrc/host/driver.cpp",
                "csrc/host/induction.cpp", "csrc/host/user_ic.cpp"]
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-result", "-Wno-unused-function"]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-pthread", "-Wall", "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _obj(src):
    return os.path.join(BUILD, os.path.basename(src).replace(".", "_") + ".o")


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        if os.path.exists(os.path.join(HERE, s)):
            jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(HERE, s), "-o", _obj(s)])
    for s in HOST_SOURCES:
        jobs.append([CXX] + HOST_FLAGS + ["-c", os.path.join(HERE, s), "-o", _obj(s)])
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for cmd, fut in [(c, ex.submit(_run, c)) for c in jobs]:
            fut.result()
            if verbose:
                print(" ".join(cmd[:1] + cmd[-3:]))
    objs = [c[-1] for c in jobs]
    lib = os.path.join(BUILD, "libspdk.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs)
    os.replace(tmp, lib)
    exe = os.path.join(BUILD, "spdk")
    _run([CXX, "-O2", os.path.join(HERE, "csrc/host/main.cpp"), "-o", exe, "-L" + BUILD, "-lspdk",
          "-Wl,-rpath,$ORIGIN"])
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
