"""Benchmark: DOF-updates/s of the spectral-difference step on one MI355X.

Workload (one "step" = one full ADER step, all Picard sweeps included):
3-D Kelvin-Helmholtz, 96^3 elements, p=3 (56.6 M active solution points),
seeded velocity perturbation, periodic box, gamma 1.4, inputs resident in
HBM. DOF-updates/s = active solution points x steps / evolution seconds (the
reference's zone-cycles/s, src/hydro_ader.hpp:263-271).

Forms timed on the same workload:
  plain_fp64     FP64 fused element-tile path (bit-identical to the reference
                 arithmetic), p predictor sweeps + corrector;
  mixed_fp32     FP32 state and contractions, FP64 fluxes / Riemann /
                 accumulation, the same sweep count;
  residual_fp32  mixed, Picard sweeps on the unconverged elements until the
                 element residual passes (zero-order guess);
  learned_fp32   residual_fp32 started from the trained per-element MLP guess
                 (the headline form: `value`).
Accuracy: every form's density after 20 steps at 32^3 against the FP64 form
(relative L1 and L-infinity).

Prints one JSON line with the headline throughput, the per-form figures, the
HBM roofline of the dominant kernel (the mixed Picard sweep; device time from
HIP events on its stream) and the CPU oracle timed on the host's cores on a
bounded sample of the same workload.
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
TOL = 1e-3        # element residual tolerance of the residual forms
WEIGHTS = os.path.join(ROOT, "spd_k_amd", "weights", "kh3d_p3.mlp")


def kh_args(n, p=P):
    return ["-i", os.path.join(ROOT, "tests", "inputs", "kelvin_helmholtz.athinput"), "job/fallback=false",
            f"mesh/p={p}", f"mesh/nx1={n}", f"mesh/nx2={n}", f"mesh/nx3={n}", "problem/perturb=random",
            "problem/amp=0.01", "hydro/gamma=1.4", "time/tlim=1.0", "output/dt=-1"]


FORMS = {
    "plain_fp64": ["job/path=fused"],
    "mixed_fp32": ["job/precision=mixed"],
    "residual_fp32": ["job/precision=mixed", "job/picard=residual", f"job/picard_tol={TOL}"],
    "learned_fp32": ["job/precision=mixed", "job/picard=residual", f"job/picard_tol={TOL}", "job/guess=mlp",
                     f"job/mlp_weights={WEIGHTS}"],
}


def mixed_sweep_bytes(p, dim, E):
    """Compulsory HBM bytes of the mixed sweeps of one fixed-count ADER step
    (each array read or written once per sweep; FP32 = 4 B, FP64 = 8 B), with
    n = T = p+1, V = 5, TPE = n^dim points, NT = n^(dim-1) face points:
      first sweep  : d_old (V TPE 4) + neighbour traces (2 dim NT V 4)
                     + D out (T V TPE 4) + traces out (2 dim NT T V 4)
      middle sweeps: D in + d_old + traces in (T slices) + D out + traces out
      corrector    : D in + d_old + traces in + U_sp in and out (V TPE 8 each)
    plus the element means (V 8) read by each sweep."""
    n = T = p + 1
    V = 5
    TPE, NT = n ** dim, n ** (dim - 1)
    d_old, D = V * TPE * 4, T * V * TPE * 4
    tr1, trT = 2 * dim * NT * V * 4, 2 * dim * NT * T * V * 4
    mu = V * 8
    first = d_old + tr1 + D + trT + mu
    middle = D + d_old + trT + D + trT + mu
    corr = D + d_old + trT + 2 * V * TPE * 8 + mu
    return float(first + (T - 2) * middle + corr) * E, T


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


def accuracy(n=32, steps=20):
    """Each form's density after `steps` steps at n^3 against the FP64 form
    (relative L1 and L-infinity)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as op
    import spd_k_amd
    out = {}
    with tempfile.TemporaryDirectory() as d:
        dirs = {}
        for name, extra in FORMS.items():
            dirs[name] = os.path.join(d, name)
            spd_k_amd.run(kh_args(n) + [f"time/nlim={steps}", "job/final_output=true"] + extra, dirs[name])
        g = op.grid(dirs["plain_fp64"])
        ref = op.load_sd(dirs["plain_fp64"], g, 0, 0)
        for name, dd in dirs.items():
            x = op.load_sd(dd, g, 0, 0)
            out[name] = {"rho_rel_l1_vs_fp64": float(np.abs(x - ref).mean() / np.abs(ref).mean()),
                         "rho_rel_linf_vs_fp64": float(np.abs(x - ref).max() / np.abs(ref).max())}
    return {"workload": f"{n}^3 KH p=3, {steps} steps", "forms": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=N_ELEM, help="elements per axis (default: the 96^3 workload)")
    ap.add_argument("--forms", default=",".join(FORMS), help="comma-separated subset of " + ",".join(FORMS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: the SD update here runs on a single GPU (multi-GPU halo exchange is a "
              "design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    import spd_k_amd
    assert torch.cuda.is_available(), "no GPU visible"
    if not os.path.exists(WEIGHTS):
        raise SystemExit(f"{WEIGHTS} is missing: train it with the driver's job/mode=train_guess (see DESIGN.md)")
    torch.cuda.synchronize()
    forms = [f for f in a.forms.split(",") if f]
    res = {}
    for name in forms:
        with tempfile.TemporaryDirectory() as d:
            r = spd_k_amd.run(kh_args(a.n) + [f"time/nlim={a.steps}", f"job/warmup={a.warmup}", "job/timing=true"]
                              + FORMS[name], d)
        torch.cuda.synchronize()
        r["value"] = r["dofs"] * r["steps"] / r["evolution_seconds"]
        res[name] = r
    head = "learned_fp32" if "learned_fp32" in res else forms[-1]
    h = res[head]
    base = res.get("plain_fp64")
    per_form = {}
    for name, r in res.items():
        per_form[name] = {"value": r["value"], "ms_per_step": r["evolution_seconds"] / r["steps"] * 1e3,
                          "vs_plain_fp64": (r["value"] / base["value"]) if base else None}
        if r.get("picard_steps"):
            per_form[name]["sweeps_per_step"] = r["picard_sweeps"] / r["picard_steps"]
            per_form[name]["fallback_steps"] = r["picard_fallbacks"]
    roof = None
    m = res.get("mixed_fp32")
    if m and m["kernel_launches"] > 0:
        E = (a.n + 2) ** 3
        bytes_step, launches = mixed_sweep_bytes(P, 3, E)
        avg_ms = m["kernel_ms_total"] / m["kernel_launches"]
        achieved = (bytes_step / launches) / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": m["kernel_name"],
                "avg_launch_ms": round(avg_ms, 3), "bytes_per_launch": int(bytes_step / launches)}
    out = {
        "metric": "DOF-updates/s on 1 MI355X, learned+FP32 vs plain FP64 at matched L1 error",
        "value": h["value"],
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": h["steps"],
        "warmup": a.warmup,
        "ms_per_step": h["evolution_seconds"] / h["steps"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"3-D Kelvin-Helmholtz {a.n}^3 elements p=3 (seeded perturbation), ADER step",
                   "dofs": h["dofs"], "form": head, "picard_tol": TOL},
        "forms": per_form,
        "vs_plain_fp64": per_form[head]["vs_plain_fp64"],
        "roofline": roof,
        "accuracy": None if a.no_accuracy else accuracy(),
        "cpu_baseline": None if a.no_cpu else cpu_baseline(),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
