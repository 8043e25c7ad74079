"""Reader of the driver's output files (the reference's output format,
src/output.cpp:11-50; the reference reads it with tests/spdk_io.py):
`{X,Y,Z}_N{N}p{p}_0.dat` hold the control-volume face coordinates of each
axis with nGH = 2 ghost faces per side, `W_cv_N{N}p{p}_{i}_0.dat` the
primitive variables at the control volumes, shape
(1, 6, Nz, Ny, Nx, nz, ny, nx) with one ghost element per side on active
axes."""
import glob
import os
import re

import numpy as np

NGH, nGH = 1, 2


def grid(outdir):
    """Elements per axis, the degree p and the interior control-volume faces."""
    g = {"faces": {}, "N": {}, "p": None}
    for d, name in enumerate("XYZ"):
        f = glob.glob(os.path.join(outdir, f"{name}_N*_0.dat"))[0]
        m = re.match(rf"{name}_N(\d+)p(\d+)_0.dat", os.path.basename(f))
        N, p = int(m.group(1)), int(m.group(2))
        xf = np.fromfile(f)
        g["N"][d] = N
        if N > 1:
            g["p"] = p
            g["faces"][d] = xf[nGH:len(xf) - nGH]
    return g


def output_indices(outdir, prefix="W_cv"):
    idx = []
    for f in glob.glob(os.path.join(outdir, f"{prefix}_*_0.dat")):
        m = re.match(rf"{prefix}_N\d+(?:p\d+)?_(\d+)_0.dat", os.path.basename(f))
        if m:
            idx.append(int(m.group(1)))
    return sorted(idx)


def load_sd(outdir, g, i, var=0, prefix="W_cv", nvar=6):
    """One variable of output i at the interior control volumes, as a
    (Nz*nz, Ny*ny, Nx*nx) array in z, y, x order."""
    n = g["p"] + 1
    act = {d: g["N"][d] > 1 for d in range(3)}
    Ne = lambda d: g["N"][d] + 2 * NGH if act[d] else 1
    npt = lambda d: n if act[d] else 1
    shape = (1, nvar, Ne(2), Ne(1), Ne(0), npt(2), npt(1), npt(0))
    A = np.fromfile(os.path.join(outdir, f"{prefix}_N{g['N'][0]}p{g['p']}_{i}_0.dat")).reshape(shape)
    v = A[0, var]
    sl = lambda d: slice(NGH, -NGH) if act[d] else slice(None)
    v = v[sl(2), sl(1), sl(0)]
    Nz, Ny, Nx, nz, ny, nx = v.shape
    return v.transpose(0, 3, 1, 4, 2, 5).reshape(Nz * nz, Ny * ny, Nx * nx)


def element_means(outdir, var=0, i=None):
    """Volume averages of one variable over each element (3-D output), from
    output i (default: the last): array (Nz, Ny, Nx)."""
    g = grid(outdir)
    n = g["p"] + 1
    x = load_sd(outdir, g, output_indices(outdir)[-1] if i is None else i, var)
    w = [np.diff(g["faces"][d]) for d in range(3)]
    vol = w[2][:, None, None] * w[1][None, :, None] * w[0][None, None, :]
    Nz, Ny, Nx = (s // n for s in x.shape)
    num = (x * vol).reshape(Nz, n, Ny, n, Nx, n).sum(axis=(1, 3, 5))
    return num / vol.reshape(Nz, n, Ny, n, Nx, n).sum(axis=(1, 3, 5))


def l1_vs_exact_sine(outdir, i, t, amp=0.1):
    """L1 of the density cell averages against the exact advected sine
    rho = 1 + amp sin(2 pi (x + y - 2t)) (2-D and 3-D runs of the sine_wave
    problem with u = v = 1; the reference's check, tests/run_tests.py:183-202)."""
    g = grid(outdir)
    xf = g["faces"][0]
    a = 2 * np.pi
    rho = load_sd(outdir, g, i, 0)
    F = lambda x, y: -np.sin(a * (x + y - 2 * t)) / a ** 2
    X0, Y0 = np.meshgrid(xf[:-1], xf[:-1], indexing="ij")
    X1, Y1 = np.meshgrid(xf[1:], xf[1:], indexing="ij")
    exact = 1.0 + amp * (F(X1, Y1) - F(X0, Y1) - F(X1, Y0) + F(X0, Y0)) / ((X1 - X0) * (Y1 - Y0))
    return float(np.abs(rho - exact.T[None, :, :]).mean())
