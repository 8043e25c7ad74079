"""Mutation check of the oracle's pins (test infrastructure, run by hand):
plants one plausible mistake at a time in a copy of oracle_solver.cpp,
builds that copy into a temporary directory, runs the CPU pin tests against
it (SPDK_ORACLE_BUILD) and reports whether a test failed. The oracle's own
sources and build are never touched.

    python oracle/mutation_check.py [mutation name ...]
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FILES = ["oracle_poly.cpp", "oracle_solver.cpp", "oracle_main.cpp", "oracle_capi.cpp", "oracle_mlp.cpp",
         "oracle_core.hpp", "Makefile"]
TESTS = ["tests/test_oracle_fallback.py", "tests/test_oracle_pins.py"]

M = [
 ("theta weights permuted", "const double w[4] = {1.0, 0.75, 0.5, 0.375};", "const double w[4] = {1.0, 0.5, 0.75, 0.375};"),
 ("ring radius 0.5", "if (ti(0, k + dk, j + dj, i + di) > 0) ring = 0.25;", "if (ti(0, k + dk, j + dj, i + di) > 0) ring = 0.5;"),
 ("fallback face range short by one", "Mx = W.Nx - 2 * cfg.fgh[0] + (cfg.fgh[0] > 0);", "Mx = W.Nx - 2 * cfg.fgh[0];"),
 ("hancock transverse sign", "dWt[D_] -= W[VY] * dWy[D_] + W[D_] * dWy[VY];", "dWt[D_] += W[VY] * dWy[D_] + W[D_] * dWy[VY];"),
 ("hancock dt scale", "WR[v] = w[v] - dw + dwt[v] * dt / h;\n        WL[v] = w[v] + dw + dwt[v] * dt / h;", "WR[v] = w[v] - dw + dwt[v] * dt;\n        WL[v] = w[v] + dw + dwt[v] * dt;"),
 ("minmod half slope dropped", "return 0.5 * slope * (xR - xL);", "return slope * (xR - xL);"),
 ("face theta uses own cell only", "double th = smax(theta(0, k, j, i), theta(0, k - (D == 2), j - (D == 1), i - (D == 0)));", "double th = theta(0, k, j, i);"),
 ("reflective FV ghost index", "BL[bidx(v, k, j, i)] = sgn * at(v, k, j, i, 2 * nGH - 1 - l);", "BL[bidx(v, k, j, i)] = sgn * at(v, k, j, i, 2 * nGH - l);"),
 ("gradfree FV ghosts mirrored", "BL[bidx(v, k, j, i)] = at(v, k, j, i, nGH + l);\n                BR[bidx(v, k, j, i)] = at(v, k, j, i, N - 2 * nGH + l);", "BL[bidx(v, k, j, i)] = at(v, k, j, i, 2 * nGH - 1 - l);\n                BR[bidx(v, k, j, i)] = at(v, k, j, i, N - 2 * nGH + l);"),
 ("gradfree SD trace point swapped", "BL = at(t, v, k, j, i, kk, jj, ii, 1, 0);\n                    BR = at(t, v, k, j, i, kk, jj, ii, N - 2, n - 1);\n                } else {", "BL = at(t, v, k, j, i, kk, jj, ii, 1, n - 1);\n                    BR = at(t, v, k, j, i, kk, jj, ii, N - 2, n - 1);\n                } else {"),
 ("reflective SD sign on wrong variable", "double sgn = (v == 1 + dim) ? -1.0 : 1.0;\n                    BL = sgn", "double sgn = (v == 2 - dim) ? -1.0 : 1.0;\n                    BL = sgn"),
 ("NAD Moore box x-only", "for (int dj = -(int)ay; dj <= (int)ay; dj++)\n                        for (int di = -(int)ax; di <= (int)ax; di++) {\n                            double u = U(v, k + dk, j + dj, i + di);", "for (int dj = 0; dj <= 0; dj++)\n                        for (int di = -(int)ax; di <= (int)ax; di++) {\n                            double u = U(v, k + dk, j + dj, i + di);"),
 ("NAD band uses tol twice", "mx += std::fabs(mx) * tol;", "mx += 2 * std::fabs(mx) * tol;"),
 ("SED range 1..N-2", "int Mz = U.Nz - 4 * cfg.ngh[2], My = U.Ny - 4 * cfg.ngh[1], Mx = U.Nx - 4 * cfg.ngh[0];\n    for3(Mz, My, Mx, 2 * cfg.ngh[2], 2 * cfg.ngh[1], 2 * cfg.ngh[0]", "int Mz = U.Nz - 2 * cfg.ngh[2], My = U.Ny - 2 * cfg.ngh[1], Mx = U.Nx - 2 * cfg.ngh[0];\n    for3(Mz, My, Mx, cfg.ngh[2], cfg.ngh[1], cfg.ngh[0]"),
 ("SED dv factor", "double dv = 0.5 * d2u * (f[l + 1] - f[l]);", "double dv = d2u * (f[l + 1] - f[l]);"),
 ("relax stencil own cell only", "if (ax) alpha = smin(smin(ax_(v, k, j, i - 1), ax_(v, k, j, i)), ax_(v, k, j, i + 1));", "if (ax) alpha = ax_(v, k, j, i);"),
 ("PAD exclusive bound", "if (d < RHO_MIN || d > RHO_MAX) tr(nvar, k, j, i) = 1;", "if (d <= RHO_MIN || d > RHO_MAX) tr(nvar, k, j, i) = 1;"),
 ("HLLC star pressure sign", "double p_star = (rc_R * W_L[P_] + rc_L * W_R[P_] + rc_L * rc_R * (v_L - v_R)) / (rc_R + rc_L);", "double p_star = (rc_R * W_L[P_] + rc_L * W_R[P_] - rc_L * rc_R * (v_L - v_R)) / (rc_R + rc_L);"),
]


def main(names):
    orig = open(os.path.join(HERE, "oracle_solver.cpp")).read()
    chosen = [m for m in M if m[0] in names] or M
    results = []
    for name, a, b in chosen:
        if a not in orig:
            results.append((name, "pattern not found"))
            continue
        with tempfile.TemporaryDirectory() as tmp:
            for f in FILES:
                shutil.copy(os.path.join(HERE, f), tmp)
            open(os.path.join(tmp, "oracle_solver.cpp"), "w").write(orig.replace(a, b, 1))
            subprocess.run(["make", "-C", tmp, "-j8"], check=True, capture_output=True)
            env = dict(os.environ, SPDK_ORACLE_BUILD=os.path.join(tmp, "_build"))
            r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider"] + TESTS,
                               cwd=ROOT, env=env, capture_output=True, text=True)
        failed = [l[7:] for l in r.stdout.splitlines() if l.startswith("FAILED")]
        results.append((name, ("caught by " + failed[0].split(" - ")[0]) if r.returncode else "MISSED"))
        print(f"{name}: {results[-1][1]}", flush=True)
    missed = [n for n, r in results if r != "caught" and not r.startswith("caught")]
    print(f"{len(results) - len(missed)} of {len(results)} mutations caught")
    return 1 if missed else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
