"""Benchmark: DOF-updates/s of the spectral-difference step on one MI355X.

Workload (one "step" = one full ADER step, all p+1 Picard sweeps included):
3-D Kelvin-Helmholtz, 96^3 elements, p=3 (56.6 M active solution points),
seeded velocity perturbation, periodic box, gamma 1.4, plain FP64 form,
inputs resident in HBM. DOF-updates/s = active solution points x steps /
evolution seconds (the reference's zone-cycles/s, src/hydro_ader.hpp:263-271).

Prints one JSON line with the throughput, the HBM roofline of the dominant
kernel (the Picard sweep; device time from HIP events on its stream) and
the CPU oracle timed on the host's cores on a bounded sample of the same
workload.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_ELEM = 96
P = 3


def kh_args(n, p=P):
    return ["-i", os.path.join(ROOT, "tests", "inputs", "kelvin_helmholtz.athinput"), "job/fallback=false",
            f"mesh/p={p}", f"mesh/nx1={n}", f"mesh/nx2={n}", f"mesh/nx3={n}", "problem/perturb=random",
            "problem/amp=0.01", "hydro/gamma=1.4", "time/tlim=1.0", "output/dt=-1"]


def sweep_bytes_per_step(p, dim, n_elem_total):
    """Compulsory HBM bytes of the Picard-sweep kernels in one ADER step of
    the fused path (reads and writes of state and face traces, each once):
    per element, with n=p+1, T=p+1, V=5 physical variables, NT=n^(dim-1)
    face points, TPE=n^dim solution points, 8-byte words:
      first sweep  : U_sp (V*TPE) + neighbour traces (2*dim*NT*V)
                     + U_ader out (T*V*TPE) + traces out (2*dim*NT*T*V)
      middle sweeps: U_ader + U_sp + traces in + U_ader out + traces out
      last sweep   : U_ader + U_sp + traces in + U_sp out
    (the trace kernel that seeds the first sweep is not a sweep)."""
    n = p + 1
    T = p + 1
    V = 5
    NT = n ** (dim - 1)
    TPE = n ** dim
    usp = V * TPE
    ua = T * V * TPE
    tr1 = 2 * dim * NT * V
    trT = T * tr1
    first = usp + tr1 + ua + trT
    middle = ua + usp + trT + ua + trT
    last = ua + usp + trT + usp
    per_elem = first + (T - 2) * middle + last
    return 8.0 * per_elem * n_elem_total, T  # bytes per step, sweep launches per step


def cpu_baseline(sample_n=24, steps=3):
    """The CPU oracle (OpenMP, every host thread OMP_NUM_THREADS allows) on a
    sample of the same workload: sample_n^3 elements, p=3, `steps` steps."""
    exe = os.path.join(ROOT, "oracle", "_build", "spdk_oracle_omp")
    if not os.path.exists(exe):
        return None
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([exe] + kh_args(sample_n) + [f"time/nlim={steps}", "-o", d], capture_output=True, text=True,
                           timeout=900, env=dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close"))
    m = re.search(r"evolution: (\d+) steps, ([0-9.eE+-]+) s .* ([0-9.eE+-]+) zone-cycles/s", r.stdout)
    if r.returncode != 0 or not m:
        return None
    return {"value": float(m.group(3)), "unit": "DOF-updates/s", "cores": cores, "kind": "port",
            "sample": f"CPU oracle (OpenMP, -O3) on the same workload at {sample_n}^3 elements p=3, {m.group(1)} ADER "
                      f"steps, evolution loop {float(m.group(2)):.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=N_ELEM, help="elements per axis (default: the 96^3 workload)")
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: the SD update here runs on a single GPU (multi-GPU halo exchange is a "
              "design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    import spd_k_amd
    assert torch.cuda.is_available(), "no GPU visible"
    torch.cuda.synchronize()
    with tempfile.TemporaryDirectory() as d:
        res = spd_k_amd.run(kh_args(a.n) + [f"time/nlim={a.steps}", f"job/warmup={a.warmup}", "job/timing=true",
                                           "job/path=fused"], d)
    torch.cuda.synchronize()
    steps = res["steps"]
    sec = res["evolution_seconds"]
    value = res["dofs"] * steps / sec
    n_tot = (a.n + 2) ** 3
    bytes_step, launches_per_step = sweep_bytes_per_step(P, 3, n_tot)
    roof = None
    if res["kernel_launches"] > 0:
        avg_ms = res["kernel_ms_total"] / res["kernel_launches"]
        achieved = (bytes_step / launches_per_step) / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": res["kernel_name"], "avg_launch_ms": round(avg_ms, 3),
                "bytes_per_launch": int(bytes_step / launches_per_step)}
    out = {
        "metric": "DOF-updates/s on 1 MI355X, learned+FP32 vs plain FP64 at matched L1 error",
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": sec / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3-D Kelvin-Helmholtz {a.n}^3 elements p=3 (seeded perturbation), ADER step, "
                               f"plain FP64 form, fused element-tile path",
                   "dofs": res["dofs"], "form": "plain-fp64"},
        "roofline": roof,
        "cpu_baseline": None if a.no_cpu else cpu_baseline(),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
