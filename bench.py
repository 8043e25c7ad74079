"""Benchmark: DOF-updates/s of the spectral-difference step on one MI355X.

Workload (one "step" = one full ADER step, all Picard sweeps and the FV
update included): 3-D Kelvin-Helmholtz, 96^3 elements, p=3 (56.6 M active
solution points), seeded velocity perturbation, periodic box, gamma 1.4,
with trouble detection (NAD + SED + PAD) and the blended MUSCL-Hancock +
HLLC fallback on, as the reference's kelvin_helmholtz input has it; inputs
resident in HBM. DOF-updates/s = active solution points x steps / evolution
seconds (the reference's zone-cycles/s, src/hydro_ader.hpp:263-271).

Forms timed on the same workload (the flux evaluation, FV update, NAD and
fallback are FP64 in every form):
  plain_fp64     FP64 throughout (bit-identical to the reference arithmetic):
                 p predictor sweeps and the flux evaluation in the element-tile
                 kernel, then the FV update with trouble detection and blending;
  mixed_fp32     the predictor with FP32 state, transforms and contractions
                 and FP64 fluxes / Riemann solves, the same sweep count;
  residual_fp32  mixed; each element sweeps until its residual passes 1e-6,
                 at most the fixed count (zero-order guess);
  learned_fp32   residual_fp32 started from the trained per-element MLP guess
                 (one sweep from it: a cap of 1 measured 110 ms/step against
                 119 with the cap at the fixed count, 96^3); elements still
                 unconverged restart from the zero-order guess with the fixed
                 count.
`value` is the learned_fp32 form, the form the metric names; the other forms
are reported beside it.

Matched L1 error, measured in the same run for every form:
  kh3d    the timed recipe (seeded perturbation on a 48^3 noise lattice, so
          both resolutions start from one field; fallback on) at 48^3, p=3,
          to t=0.03: L1 of the element means of density against the FP64 run
          at 96^3 (averaged onto the 48^3 elements);
  sine3d  3-D sine advection, 24^3, p=3, t=0.2: L1 of density against the
          exact solution.
A form matches when both errors are within 1.1x of the plain FP64 form's.

Also measured (the "induction" object): the 2-D field-loop induction
configuration, 1024^2 elements p=3, stage sequence vs fused FP64 vs mixed,
with its one-period matched-error check.

Prints one JSON line with the headline throughput, the per-form figures, the
HBM roofline of the dominant kernel (device time from HIP events on its
stream, HBM traffic from rocprofv3 FETCH_SIZE / WRITE_SIZE scaled by the
same counters on a device copy of known size) and the CPU oracle timed on the
host's cores on a bounded sample of the same workload.
"""
import argparse
import csv
import glob
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_ELEM = 96
P = 3
TOL = 1e-6        # element residual tolerance of the residual forms
MATCH = 1.1       # matched L1 error: within this factor of the FP64 form's
WEIGHTS = os.path.join(ROOT, "spd_k_amd", "weights", "kh3d_p3.mlp")
WEIGHTS_2D_P4 = os.path.join(ROOT, "spd_k_amd", "weights", "kh2d_p4.mlp")  # cfg4's network (106 -> 256 -> 256 -> 500)
DRIVER = os.path.join(ROOT, "spd_k_amd", "_build", "spdk")


HEADLINE = "learned_fp32"
NOISE_LATTICE = 48  # accuracy runs: one noise sample per cell of a 48^3 lattice (divides 48 and 96)


def kh_args(n, p=P, noise_cells=0):
    """The timed recipe: 3-D KH with the seeded perturbation and the FV
    fallback (the input file's job/fallback=true)."""
    a = ["-i", os.path.join(ROOT, "tests", "inputs", "kelvin_helmholtz.athinput"), "job/fallback=true",
         f"mesh/p={p}", f"mesh/nx1={n}", f"mesh/nx2={n}", f"mesh/nx3={n}", "problem/amp=0.01", "hydro/gamma=1.4",
         "problem/perturb=random", "time/tlim=1.0", "output/dt=-1"]
    if noise_cells:
        a.append(f"problem/noise_cells={noise_cells}")
    return a


FORMS = {
    "plain_fp64": ["job/path=fused"],
    "mixed_fp32": ["job/precision=mixed"],
    "residual_fp32": ["job/precision=mixed", "job/picard=residual", f"job/picard_tol={TOL}"],
    "learned_fp32": ["job/precision=mixed", "job/picard=residual", f"job/picard_tol={TOL}", "job/guess=mlp",
                     f"job/mlp_weights={WEIGHTS}", "job/picard_max=1"],
}


def mixed_sweep_bytes(p, dim, E, corrector=False):
    """Compulsory HBM bytes of the mixed predictor sweeps of one fixed-count
    ADER step, and of the corrector when it runs in the same kernel (each
    array read or written once per sweep; FP32 = 4 B, FP64 = 8 B), with
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
    if corrector:
        return float(first + (T - 2) * middle + corr) * E, T
    return float(first + (T - 2) * middle) * E, T - 1


def face_sweep_bytes(p, dim, E, mixed):
    """Compulsory HBM bytes of the flux-evaluation sweep that forms the FV face
    integrals of every ADER node (the fallback step's last sweep): the
    predictor slices in (T V TPE; FP32 deviations in the mixed form, FP64
    otherwise), the lower neighbour's face trace per axis (each element solves
    the Riemann problem at its lower faces only: dim NT T V, same type), the
    element means (mixed: V 8), and the face arrays out: per element and axis
    its TPE lower faces, V variables, T nodes, FP64 (T dim V TPE 8)."""
    n = T = p + 1
    V = 5
    TPE, NT = n ** dim, n ** (dim - 1)
    w = 4 if mixed else 8
    return float(T * V * TPE * w + dim * NT * T * V * w + (V * 8 if mixed else 0) + T * dim * V * TPE * 8) * E


def fp64_sweep_bytes(p, dim, E):
    """Compulsory HBM bytes of the p predictor sweeps of the FP64 element-tile
    kernel in one ADER step (each array once per sweep, FP64):
      first sweep  : U_sp (V TPE) + neighbour traces (2 dim NT V) + slices
                     out (T V TPE) + traces out (2 dim NT T V)
      later sweeps : slices in + U_sp + traces in (T slices) + slices out +
                     traces out.
    Returns (bytes, sweeps)."""
    n = T = p + 1
    V = 5
    TPE, NT = n ** dim, n ** (dim - 1)
    U, S = V * TPE * 8, T * V * TPE * 8
    tr1, trT = 2 * dim * NT * V * 8, 2 * dim * NT * T * V * 8
    first = U + tr1 + S + trT
    later = S + U + trT + S + trT
    return float(first + (T - 2) * later) * E, T - 1


READ_REQ = ("TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum", "TCC_EA0_RDREQ_sum")


def hbm_traffic(n, steps=1):
    """HBM bytes per predictor-sweep launch of the fixed-count mixed step, from two
    rocprofv3 --pmc passes: the L2's memory read requests by size (32, 64 and
    128 bytes, and their total), and WRITE_SIZE (KiB per dispatch; write
    requests are at most 64 bytes, which it counts by size). FETCH_SIZE is not
    used: it counts a 128-byte read request as 64 bytes. The warm-up's device
    copy of U_sp, a plain copy of known size, checks both (copy_check: measured
    / copied bytes, reads and writes). None when rocprofv3 is absent or fails."""
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None
    E = (n + 2) ** 3
    copy_bytes = 6 * E * (P + 1) ** 3 * 8
    per = {}  # counter -> (sweep average, largest copy reading)
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        for name, ctrs in (("reads", READ_REQ), ("writes", ("WRITE_SIZE",))):
            out = os.path.join(d, name)
            cmd = [prof, "--pmc"] + list(ctrs) + ["--output-format", "csv", "-d", out, "-o", "p", "--", DRIVER] + \
                kh_args(n) + [f"time/nlim={steps}", "job/warmup=1", "job/precision=mixed"]
            try:
                r = subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), capture_output=True,
                                   text=True, timeout=180)
            except subprocess.TimeoutExpired:
                return None
            files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
            if r.returncode != 0 or not files:
                return None
            acc = {c: ({}, {}) for c in ctrs}  # counter -> (copy dispatches, sweep dispatches)
            for row in csv.DictReader(open(files[0])):
                c = row["Counter_Name"]
                if c not in acc:
                    continue
                k = row["Kernel_Name"]
                slot = 0 if "copyBuffer" in k else (1 if "k_msweep<4, 3, 0>" in k else None)
                if slot is not None:
                    dd = acc[c][slot]
                    dd[row["Dispatch_Id"]] = dd.get(row["Dispatch_Id"], 0.0) + float(row["Counter_Value"])
            for c, (cp, sw) in acc.items():
                if not cp or not sw:
                    return None
                per[c] = (sum(sw.values()) / len(sw), max(cp.values()))
    size = {"TCC_EA0_RDREQ_32B_sum": 32.0, "TCC_EA0_RDREQ_64B_sum": 64.0, "TCC_EA0_RDREQ_128B_sum": 128.0}
    reads = sum(per[c][0] * w for c, w in size.items())
    reads_copy = sum(per[c][1] * w for c, w in size.items())
    sized = sum(per[c][0] for c in size)  # requests of a known size, against the total
    writes, writes_copy = per["WRITE_SIZE"][0] * 1024.0, per["WRITE_SIZE"][1] * 1024.0
    return {"bytes_per_launch": reads + writes, "fetch_bytes": reads, "write_bytes": writes,
            "read_requests": {c.split("_")[3]: per[c][0] for c in size},
            "read_requests_sized_fraction": sized / per["TCC_EA0_RDREQ_sum"][0] if per["TCC_EA0_RDREQ_sum"][0] else None,
            "copy_check": {"reads": reads_copy / copy_bytes, "writes": writes_copy / copy_bytes}}


def cpu_baseline(sample_n=24, steps=12):
    """The CPU oracle (OpenMP, every host thread OMP_NUM_THREADS allows) on a
    sample of the same workload: sample_n^3 elements, p=3, `steps` steps (about
    10 s of CPU work on 16 threads)."""
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


def accuracy(forms, coarse=48, t_end=0.03):
    """Matched-L1 check of every form (module docstring)."""
    import numpy as np
    import spd_k_amd
    from spd_k_amd.outputs import element_means, l1_vs_exact_sine, output_indices
    sine = ["-i", os.path.join(ROOT, "tests", "inputs", "sine_wave.athinput"), "mesh/nx1=24", "mesh/nx2=24",
            "mesh/nx3=24", "mesh/p=3", "problem/amp=0.1", "time/tlim=0.2", "output/dt=0.2", "output/fields=W_cv"]
    kh = lambda n: kh_args(n, noise_cells=NOISE_LATTICE)[:-2] + [f"time/tlim={t_end}", f"output/dt={t_end}",
                                                                 "output/fields=W_cv"]
    out = {}
    with tempfile.TemporaryDirectory() as d:
        fine = os.path.join(d, "kh_fine")
        spd_k_amd.run(kh(2 * coarse) + FORMS["plain_fp64"], fine)
        ref = element_means(fine)
        shutil.rmtree(fine, ignore_errors=True)
        ref = ref.reshape(ref.shape[0] // 2, 2, ref.shape[1] // 2, 2, ref.shape[2] // 2, 2).mean(axis=(1, 3, 5))
        for name in forms:
            sd, kd = os.path.join(d, "s_" + name), os.path.join(d, "k_" + name)
            spd_k_amd.run(sine + FORMS[name], sd)
            spd_k_amd.run(kh(coarse) + FORMS[name], kd)
            out[name] = {"sine3d_l1": l1_vs_exact_sine(sd, output_indices(sd)[-1], 0.2, amp=0.1),
                         "kh3d_l1": float(np.abs(element_means(kd) - ref).mean() / np.abs(ref).mean())}
    base = out.get("plain_fp64")
    for name, r in out.items():
        if base:
            r["sine3d_vs_fp64"] = r["sine3d_l1"] / base["sine3d_l1"]
            r["kh3d_vs_fp64"] = r["kh3d_l1"] / base["kh3d_l1"]
            r["matched"] = r["sine3d_vs_fp64"] <= MATCH and r["kh3d_vs_fp64"] <= MATCH
    return {"sine3d": "3-D sine advection 24^3 p=3 to t=0.2, density L1 vs exact",
            "kh3d": f"3-D KH (timed recipe, fallback on, {NOISE_LATTICE}^3 noise lattice) {coarse}^3 p=3 to "
                    f"t={t_end}, element-mean density L1 vs FP64 at {2 * coarse}^3",
            "match_factor": MATCH, "forms": out}


def induction_loop(n=1024, steps=6, warmup=2):
    """The 2-D field-loop induction configuration: n^2 elements, p=3,
    v = (1, 0.5), A_z = 1e-3 (0.3 - r) inside r < 0.3, periodic [0,2]^2 (the
    [-1,1]^2 box shifted by (1,1)). DOF-updates/s of the stage sequence, the
    fused FP64 step and the mixed step; and the matched-error check: after
    one period (t = 4) the exact field is the initial one; L1 of the B cell
    averages against it at 48^2 for the FP64 and the mixed form."""
    import numpy as np
    import spd_k_amd
    from spd_k_amd.outputs import grid, load_sd
    base = ["-i", os.path.join(ROOT, "tests", "inputs", "induction_loop.athinput"), "job/fallback=false",
            "mesh/nx3=1", "mesh/x1len=2", "mesh/x2len=2", "induction/v1=1.0", "induction/v2=0.5", "induction/v3=0.0",
            "induction/cx=1.0", "induction/cy=1.0", "induction/nu=0"]
    forms = {"stages_fp64": ["job/path=stages"], "fused_fp64": [], "fused_mixed": ["job/precision=mixed"]}
    out = {"workload": f"2-D field-loop induction {n}^2 elements p=3, v=(1,0.5), ADER step", "forms": {}}
    with tempfile.TemporaryDirectory() as d:
        for name, extra in forms.items():
            r = spd_k_amd.run(base + [f"mesh/nx1={n}", f"mesh/nx2={n}", "time/tlim=8", f"time/nlim={steps}",
                                      f"job/warmup={warmup}", "output/dt=-1"] + extra, os.path.join(d, name))
            out["forms"][name] = {"value": r["dofs"] * r["steps"] / r["evolution_seconds"],
                                  "ms_per_step": r["evolution_seconds"] / r["steps"] * 1e3}
        l1 = {}
        for name in ("fused_fp64", "fused_mixed"):
            od = os.path.join(d, "period_" + name)
            spd_k_amd.run(base + ["mesh/nx1=48", "mesh/nx2=48", "time/tlim=4", "output/dt=4"] + forms[name], od)
            g = grid(od)
            b0 = np.stack([load_sd(od, g, 0, v, "B2_cv", 4) for v in (0, 1)])
            b1 = np.stack([load_sd(od, g, 1, v, "B2_cv", 4) for v in (0, 1)])
            l1[name] = float(np.abs(b1 - b0).mean())
    ref = out["forms"]["fused_fp64"]["value"]
    for f in out["forms"].values():
        f["vs_fused_fp64"] = f["value"] / ref
    out["one_period_l1_48"] = l1
    out["mixed_l1_vs_fp64"] = l1["fused_mixed"] / l1["fused_fp64"]
    return out


def all_configs(steps=6, warmup=2):
    """DOF-updates/s of every synthetic configuration on one GPU, each form
    the driver offers for it (`--all-configs`; the default run measures the
    3-D KH workload and the 2-D induction one):
      cfg1 2-D sine advection 32^2 p=3, ADER, one period (t = 1), L1 of
           density against the exact solution;
      cfg2 Sod (256^2, x gradfree, y periodic) and the central-cells Sedov
           blast (256^2, P = 1e-5 background), p=3, NAD + MUSCL fallback: the
           stage sequence and the default (fused predictor and flux sweep,
           stage FV update), `steps` steps;
      cfg4 2-D KH 512^2 p=4 (seeded perturbation, CFL 0.5 without the
           fallback), FP64 / mixed / residual (1e-6 and 1e-10) forms, and the
           residual forms from the learned guess (spd_k_amd/weights/kh2d_p4.mlp)
           with their sweep histograms.
    cfg3 and cfg5 are the default run's workloads."""
    import spd_k_amd
    from spd_k_amd.outputs import l1_vs_exact_sine, output_indices
    inp = lambda f: os.path.join(ROOT, "tests", "inputs", f)
    out = {}

    def timed(args, d, full=False):
        extra = [] if full else [f"time/nlim={steps}", f"job/warmup={warmup}", "output/dt=-1"]
        r = spd_k_amd.run(args + extra, d)
        rec = {"value": r["dofs"] * r["steps"] / r["evolution_seconds"],
               "ms_per_step": r["evolution_seconds"] / r["steps"] * 1e3, "steps": r["steps"], "dofs": r["dofs"]}
        if r.get("picard_steps"):
            rec["sweeps_per_step"] = r["picard_sweeps"] / r["picard_steps"]
            h = r.get("picard_hist") or []
            if any(h):  # elements by predictor sweeps to the tolerance (bin 31: capped)
                rec["histogram"] = {("capped" if b == 31 else str(b)): int(n) for b, n in enumerate(h) if n}
        return rec

    with tempfile.TemporaryDirectory() as d:
        sine = ["-i", inp("sine_wave.athinput"), "mesh/nx1=32", "mesh/nx2=32", "mesh/nx3=1", "problem/amp=0.1",
                "time/tlim=1.0", "output/dt=1.0"]
        c1 = {}
        for name, extra in (("plain_fp64", FORMS["plain_fp64"]), ("mixed_fp32", FORMS["mixed_fp32"]),
                            ("residual_fp32", FORMS["residual_fp32"])):
            od = os.path.join(d, "c1_" + name)
            c1[name] = timed(sine + extra, od, full=True)
            c1[name]["l1_vs_exact"] = l1_vs_exact_sine(od, output_indices(od)[-1], 1.0, amp=0.1)
        out["cfg1_sine2d_32_p3"] = c1
        sod = ["-i", inp("sod.athinput"), "mesh/nx1=256", "mesh/nx2=256", "mesh/x2_bc=periodic", "time/tlim=0.2"]
        sedov = ["-i", inp("sine_wave.athinput"), "problem/problem=sedov_cells", "job/fallback=true", "mesh/nx1=256",
                 "mesh/nx2=256", "mesh/nx3=1", "hydro/gamma=1.4", "time/tlim=0.05"]
        for key, args in (("cfg2_sod_256_p3_fallback", sod), ("cfg2_sedov_256_p3_fallback", sedov)):
            out[key] = {"fp64_stages": timed(args + ["job/path=stages"], os.path.join(d, key + "_s")),
                        "fp64_fused": timed(args, os.path.join(d, key + "_f"))}
        kh2 = ["-i", inp("kelvin_helmholtz.athinput"), "job/fallback=false", "mesh/nx1=512", "mesh/nx2=512",
               "mesh/nx3=1", "mesh/p=4", "problem/perturb=random", "problem/amp=0.01", "time/cfl=0.5", "time/tlim=1"]
        c4 = {}
        for name, extra in (("plain_fp64", FORMS["plain_fp64"]), ("mixed_fp32", FORMS["mixed_fp32"]),
                            ("residual_fp32", FORMS["residual_fp32"]),
                            ("residual_fp32_tol1e-10", ["job/precision=mixed", "job/picard=residual",
                                                        "job/picard_tol=1e-10"])):
            c4[name] = timed(kh2 + extra, os.path.join(d, "c4_" + name))
        if os.path.exists(WEIGHTS_2D_P4):
            for tol in ("1e-6", "1e-10"):
                c4[f"learned_fp32_tol{tol}"] = timed(kh2 + ["job/precision=mixed", "job/picard=residual",
                                                            f"job/picard_tol={tol}", "job/guess=mlp",
                                                            f"job/mlp_weights={WEIGHTS_2D_P4}"],
                                                     os.path.join(d, "c4_learned" + tol))
        out["cfg4_kh2d_512_p4"] = c4
    return out


def zone_cycles(integrator="rk3", R_list=(32, 64, 128, 256), p_list=(3, 7), sims=("sd", "sdfb")):
    """The reference's zone-cycles/s sweep (tests/benchmark.py:32-76): 3-D
    advected sine wave, RK3, without (sd) and with (sdfb) the FV fallback,
    p in {3, 7}, R = N (p+1) cells per axis, tlim = 1.4 / R (about 10 steps),
    outputs off, the driver's evolution clock. One row per configuration, in
    the reference's row format, with the form this driver runs (the fused
    tile kernel where a variant exists, else the stage sequence)."""
    import spd_k_amd
    rows = []
    with tempfile.TemporaryDirectory() as d:
        for sim in sims:
            for p in p_list:
                for R in R_list:
                    N = R // (p + 1)
                    args = ["-i", os.path.join(ROOT, "tests", "inputs", "sine_wave.athinput"), f"mesh/p={p}",
                            f"mesh/nx1={N}", f"mesh/nx2={N}", f"mesh/nx3={N}",
                            f"job/fallback={'true' if sim == 'sdfb' else 'false'}", f"time/integrator={integrator}",
                            f"time/tlim={1.4 / R}", "output/dt=-1"]
                    try:
                        r = spd_k_amd.run(args, os.path.join(d, f"{sim}_{p}_{R}"))
                        rows.append({"repo": "spd_k_amd", "sim": sim, "p": p, "R": R, "steps": r["steps"],
                                     "ms_per_step": r["evolution_seconds"] / max(r["steps"], 1) * 1e3,
                                     "zcps": r["dof_updates_per_s"]})
                    except RuntimeError as e:
                        rows.append({"repo": "spd_k_amd", "sim": sim, "p": p, "R": R, "error": str(e)[-300:]})
    return {"zone_cycles": rows, "protocol": "tests/benchmark.py of the reference: 3-D sine, " + integrator +
            ", sd and sdfb, p in {3,7}, R = N (p+1) in {32,64,128,256}, tlim = 1.4/R, outputs off"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=N_ELEM, help="elements per axis (default: the 96^3 workload)")
    ap.add_argument("--forms", default=",".join(FORMS), help="comma-separated subset of " + ",".join(FORMS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-traffic", action="store_true")
    ap.add_argument("--no-induction", action="store_true", help="skip the 2-D induction configuration")
    ap.add_argument("--all-configs", action="store_true",
                    help="print one JSON object with the other synthetic configurations instead")
    ap.add_argument("--zone-cycles", action="store_true",
                    help="print the reference's zone-cycles/s sweep (3-D sine, RK3, sd/sdfb, p 3 and 7) instead")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: the SD update here runs on a single GPU (multi-GPU halo exchange is a "
              "design note only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    # the driver library prints its run log on fd 1: send it to stderr so
    # that stdout carries the JSON line alone
    sys.stdout.flush()
    json_fd = os.dup(1)
    os.dup2(2, 1)
    import torch
    import spd_k_amd
    assert torch.cuda.is_available(), "no GPU visible"
    forms = [f for f in a.forms.split(",") if f]
    if "learned_fp32" in forms and not os.path.exists(WEIGHTS):
        raise SystemExit(f"{WEIGHTS} is missing: train it with the driver's job/mode=train_guess")
    torch.cuda.synchronize()
    if a.all_configs or a.zone_cycles:
        res = all_configs(a.steps, a.warmup) if a.all_configs else zone_cycles()
        sys.stdout.flush()
        os.dup2(json_fd, 1)
        print(json.dumps(res), flush=True)
        return
    res = {}
    for name in forms:
        with tempfile.TemporaryDirectory() as d:
            r = spd_k_amd.run(kh_args(a.n) + [f"time/nlim={a.steps}", f"job/warmup={a.warmup}", "job/timing=true"]
                              + FORMS[name], d)
        torch.cuda.synchronize()
        r["value"] = r["dofs"] * r["steps"] / r["evolution_seconds"]
        res[name] = r
    acc = None if a.no_accuracy else accuracy(forms)
    base = res.get("plain_fp64")
    per_form = {}
    for name, r in res.items():
        per_form[name] = {"value": r["value"], "ms_per_step": r["evolution_seconds"] / r["steps"] * 1e3,
                          "vs_plain_fp64": (r["value"] / base["value"]) if base else None}
        if r.get("picard_steps"):
            per_form[name]["sweeps_per_step"] = r["picard_sweeps"] / r["picard_steps"]
            per_form[name]["steps_with_capped_elements"] = r["picard_fallbacks"]
        if acc and name in acc["forms"]:
            per_form[name]["matched_l1"] = acc["forms"][name].get("matched")
    # headline: the form the metric names (learned guess + FP32 contractions)
    head = HEADLINE if HEADLINE in res else forms[-1]
    h = res[head]
    roof = None
    m = res.get("mixed_fp32")
    if m and m["kernel_launches"] > 0:
        E = (a.n + 2) ** 3
        bytes_step, launches = mixed_sweep_bytes(P, 3, E)
        avg_ms = m["kernel_ms_total"] / m["kernel_launches"]
        achieved = (bytes_step / launches) / (avg_ms * 1e-3) / 1e9
        tr = None if a.no_traffic else hbm_traffic(a.n)
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": int(tr["bytes_per_launch"]) if tr else None,
                "kernel": "k_msweep (mixed Picard predictor sweep, fixed-count step)", "avg_launch_ms": round(avg_ms, 3),
                "bytes_per_launch": int(bytes_step / launches), "traffic_detail": tr}
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
        "config": {"workload": f"3-D Kelvin-Helmholtz {a.n}^3 elements p=3 (seeded perturbation), ADER step with "
                               "NAD + MUSCL-HLLC fallback on (FP64 FV update)",
                   "dofs": h["dofs"], "form": head, "picard_tol": TOL,
                   "matched_l1": per_form[head].get("matched_l1")},
        "vs_plain_fp64": per_form[head]["vs_plain_fp64"],
        "forms": per_form,
        "accuracy": acc,
        "roofline": roof,
        "cpu_baseline": None if a.no_cpu else cpu_baseline(),
        "induction": None if a.no_induction else induction_loop(),
    }
    sys.stdout.flush()
    os.dup2(json_fd, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
