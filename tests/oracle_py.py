"""Test helpers: the CPU oracle (oracle/_build) through ctypes and as a
command-line driver, the product driver, and output-file comparison.

Only tests/ use this module; the oracle is the checker, never the product.
"""
import ctypes
import glob
import os
import re
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# SPDK_ORACLE_BUILD points the checker at another build of the oracle (oracle/mutation_check.py)
ORACLE_DIR = os.environ.get("SPDK_ORACLE_BUILD", os.path.join(ROOT, "oracle", "_build"))
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
ORACLE_EXE = os.path.join(ORACLE_DIR, "spdk_oracle_omp")
ORACLE_EXE_SERIAL = os.path.join(ORACLE_DIR, "spdk_oracle")
DRIVER_EXE = os.path.join(ROOT, "spd_k_amd", "_build", "spdk")
INPUTS = os.path.join(ROOT, "tests", "inputs")
GOLDEN = os.path.join(ROOT, "tests", "golden")

NGH, nGH = 1, 2


class OrcCfg(ctypes.Structure):
    _fields_ = [("act", ctypes.c_int * 3), ("bc", ctypes.c_int * 3), ("gamma", ctypes.c_double),
                ("g", ctypes.c_double * 3), ("nad_tol", ctypes.c_double), ("nad_delta", ctypes.c_int),
                ("nad_moore", ctypes.c_int), ("sed", ctypes.c_int), ("blending", ctypes.c_int), ("p", ctypes.c_int)]

    @classmethod
    def make(cls, act=(1, 1, 1), bc=(0, 0, 0), gamma=1.4, g=(0, 0, 0), nad_tol=1e-5, nad_delta=0, nad_moore=1, sed=1,
             blending=1, p=3):
        c = cls()
        for d in range(3):
            c.act[d], c.bc[d], c.g[d] = int(act[d]), int(bc[d]), float(g[d])
        c.gamma, c.nad_tol, c.nad_delta, c.nad_moore, c.sed, c.blending, c.p = gamma, nad_tol, nad_delta, nad_moore, sed, blending, p
        return c


_orc = None


def orc():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError("oracle not built: make -C oracle")
        L = ctypes.CDLL(ORACLE_LIB)
        for f in ("orc_minmod", "orc_alpha", "orc_normal"):
            getattr(L, f).restype = ctypes.c_double
        L.orc_minmod.argtypes = [ctypes.c_double] * 4
        L.orc_alpha.argtypes = [ctypes.c_double] * 4
        L.orc_normal.argtypes = [ctypes.c_ulonglong, ctypes.c_ulonglong]
        L.orc_dt.restype = ctypes.c_double
        L.orc_mlp_grad.restype = ctypes.c_double
        L.orc_bf16_round.restype = ctypes.c_double
        L.orc_bf16_round.argtypes = [ctypes.c_double]
        _orc = L
    return _orc


def dp(a):
    """ctypes double* of a contiguous float64 numpy array (kept alive by the caller)."""
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def ip(shape):
    return (ctypes.c_int * len(shape))(*[int(x) for x in shape])


def matrices(p):
    """Operator set of degree p from the oracle (ADER, T = p+1)."""
    n, m, T = p + 1, p + 2, p + 1
    out = {k: np.zeros(s) for k, s in [("x_sp", n), ("x_fp", m), ("xt", T), ("wt", T), ("sp_to_fp", (m, n)),
                                       ("dfp_to_sp", (n, m)), ("sp_to_cv", (n, n)), ("cv_to_sp", (n, n)),
                                       ("fp_to_cv", (n, m)), ("ader", (T, T)), ("invader", (T, T))]}
    orc().orc_matrices(p, *[dp(out[k]) for k in ("x_sp", "x_fp", "xt", "wt", "sp_to_fp", "dfp_to_sp", "sp_to_cv",
                                                "cv_to_sp", "fp_to_cv", "ader", "invader")])
    # fp_to_sp is the Lagrange interpolation fp -> sp (not needed by the kernels' tests)
    out["fp_to_sp"] = np.zeros((n, m))
    return out


def fv_geometry(N, p, L=1.0):
    """FV sub-grid faces/centres of one active axis (src/structs.hpp:40-75)."""
    mats = matrices(p)
    x_fp = mats["x_fp"]
    h = L / N
    nsp, nfp = p + 1, p + 2
    Ntot = N + 2
    fvn = N * nsp + 4
    idL = nsp - 2
    idR = fvn + idL
    faces = np.zeros(fvn + 1)
    centers = np.zeros(fvn)
    for j in range(Ntot):
        sdf = [(j - 1 + x_fp[i]) * h for i in range(nfp)]
        for i in range(nfp):
            g = i + j * nsp
            if idL <= g < idR:
                faces[g - idL] = sdf[i]
        for i in range(nsp):
            g = i + j * nsp
            if idL <= g < idR:
                centers[g - idL] = 0.5 * (sdf[i + 1] + sdf[i])
    return faces, centers


def _oracle_env():
    """OpenMP threads for one oracle process: the host's share per pytest
    worker (spinning OpenMP barriers on oversubscribed cores make small runs
    take minutes), capped by OMP_NUM_THREADS; idle threads sleep."""
    share = max(1, (os.cpu_count() or 1) // int(os.environ.get("PYTEST_XDIST_WORKER_COUNT", "1")))
    cap = int(os.environ.get("OMP_NUM_THREADS", share))
    return dict(os.environ, OMP_NUM_THREADS=str(max(1, min(share, cap))), OMP_WAIT_POLICY="PASSIVE")


def run_oracle(args, outdir, serial=False):
    exe = ORACLE_EXE_SERIAL if serial else ORACLE_EXE
    r = subprocess.run([exe] + list(args) + ["-o", outdir], capture_output=True, text=True, timeout=600,
                       env=_oracle_env())
    if r.returncode != 0:
        raise RuntimeError(f"oracle failed: {r.stdout[-2000:]}{r.stderr[-2000:]}")
    return r.stdout


def run_driver(args, outdir, timeout=300):
    env = dict(os.environ, SPD_OUTPUT_DIR=outdir)
    r = subprocess.run([DRIVER_EXE] + list(args), capture_output=True, text=True, timeout=timeout, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"driver failed: {r.stdout[-2000:]}{r.stderr[-2000:]}")
    return r.stdout


def compare_dirs(a, b):
    """compare_outputs.py semantics plus a bit count: returns {name: (n_differ, max_rel)}."""
    fa = {os.path.basename(f) for f in glob.glob(os.path.join(a, "*.dat"))}
    fb = {os.path.basename(f) for f in glob.glob(os.path.join(b, "*.dat"))}
    assert fa == fb, f"file sets differ: {sorted(fa ^ fb)}"
    res = {}
    for n in sorted(fa):
        x = np.fromfile(os.path.join(a, n))
        y = np.fromfile(os.path.join(b, n))
        assert x.shape == y.shape, n
        same = (x == y) | (np.isnan(x) & np.isnan(y))
        scale = max(np.nanmax(np.abs(x)) if x.size else 0.0, 1e-300)
        res[n] = (int((~same).sum()), float(np.nanmax(np.abs(x - y)) / scale) if x.size else 0.0)
    return res


# ---- readers of the output format (shared with the package) ----------------
from spd_k_amd.outputs import element_means, grid, load_sd, output_indices  # noqa: E402,F401


def total_mass(outdir, i, prefix="W_cv"):
    g = grid(outdir)
    rho = load_sd(outdir, g, i, 0, prefix)
    V = np.ones(rho.shape)
    if 0 in g["faces"]:
        V = V * np.diff(g["faces"][0])[None, None, :]
    if 1 in g["faces"]:
        V = V * np.diff(g["faces"][1])[None, :, None]
    if 2 in g["faces"]:
        V = V * np.diff(g["faces"][2])[:, None, None]
    return float((rho * V).sum())


def l1_vs_exact_sine(outdir, i, t, ndim, amp=0.125):
    """L1 of rho cell averages vs the exact advected sine (run_tests.py:183-202)."""
    g = grid(outdir)
    xf = g["faces"][0]
    a = 2 * np.pi
    rho = load_sd(outdir, g, i, 0)
    if ndim == 1:
        F = lambda x: -np.cos(a * (x - t)) / a
        exact = 1.0 + amp * (F(xf[1:]) - F(xf[:-1])) / np.diff(xf)
        return float(np.abs(rho[0, 0] - exact).mean())
    F = lambda x, y: -np.sin(a * (x + y - 2 * t)) / a ** 2
    X0, Y0 = np.meshgrid(xf[:-1], xf[:-1], indexing="ij")
    X1, Y1 = np.meshgrid(xf[1:], xf[1:], indexing="ij")
    exact = 1.0 + amp * (F(X1, Y1) - F(X0, Y1) - F(X1, Y0) + F(X0, Y0)) / ((X1 - X0) * (Y1 - Y0))
    return float(np.abs(rho - exact.T[None, :, :]).mean())


def mass_drift(outdir):
    m = [total_mass(outdir, i) for i in output_indices(outdir)]
    return max(abs(x - m[0]) / abs(m[0]) for x in m)


# ---- MLP oracle ------------------------------------------------------------
def mlp_param_count(dims):
    return sum(dims[l + 1] * dims[l] + dims[l + 1] for l in range(3))


def mlp_init(dims, seed=0, scale=1.0):
    """Seeded parameters (He-like scaling) in the spdk_mlp_set_params layout."""
    rng = np.random.default_rng(seed)
    out = []
    for l in range(3):
        out.append(rng.standard_normal(dims[l + 1] * dims[l]) * scale * np.sqrt(2.0 / dims[l]))
        out.append(rng.standard_normal(dims[l + 1]) * 0.1 * scale)
    return np.concatenate(out).astype(np.float32)


def orc_mlp_forward(dims, params, X, bf16):
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.zeros((X.shape[0], dims[3]))
    orc().orc_mlp_forward(ip(dims), params.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), dp(X), X.shape[0],
                          int(bf16), dp(Y))
    return Y


def orc_mlp_grad(dims, params, X, T, bf16):
    X = np.ascontiguousarray(X, dtype=np.float64)
    T = np.ascontiguousarray(T, dtype=np.float64)
    g = np.zeros(mlp_param_count(dims))
    loss = orc().orc_mlp_grad(ip(dims), params.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), dp(X), dp(T),
                              X.shape[0], int(bf16), dp(g))
    return g, loss
