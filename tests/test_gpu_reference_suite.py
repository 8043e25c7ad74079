"""The reference's own regression suite (tests/run_tests.py:44-143 in the
reference: ten configurations, each with its checks) run through the product
driver on the GPU, with the suite's own limits: L1 of the density against the
exact advected sine, mass drift below 1e-12 over the outputs, and the golden
files at their rtol. Each run is also compared bit for bit with the CPU
oracle on the same inputs."""
import os

import numpy as np
import pytest

import oracle_py as op

pytestmark = pytest.mark.gpu

SINE = "sine_wave.athinput"
SUITE = {
    # name: (input, overrides, ndim, checks, field, t_end, l1_limit, golden_name, golden_rtol)
    "hydro_sd_3d": (SINE, [], 3, ["analytic", "mass_strict", "golden"], "W_cv_N8p3_1_0.dat", 0.1, 3.0e-5,
                    "hydro_sd", 1e-6),
    "hydro_sd_2d": (SINE, ["mesh/nx3=1"], 2, ["analytic", "mass_strict"], "W_cv_N8p3_1_0.dat", 0.1, 3.0e-5, None, None),
    "hydro_sd_1d": (SINE, ["mesh/nx2=1", "mesh/nx3=1"], 1, ["analytic", "mass_strict"], "W_cv_N8p3_1_0.dat", 0.1,
                    2.0e-5, None, None),
    "hydro_fv_3d": (SINE, ["job/fallback=true", "output/dt=0.05"], 3, ["mass_strict"], "W_cv_N8p3_2_0.dat", 0.1, None,
                    None, None),
    "hydro_fv_2d": (SINE, ["job/fallback=true", "mesh/nx3=1"], 2, ["mass_strict"], "W_cv_N8p3_1_0.dat", 0.1, None,
                    None, None),
    "hydro_fv_blast_2d": (SINE, ["job/fallback=true", "mesh/nx3=1", "problem/problem=spherical_blast",
                                 "hydro/gamma=1.6666666666667", "time/tlim=0.02", "output/dt=0.02"], 2,
                          ["mass_strict"], "W_cv_N8p3_1_0.dat", 0.02, None, None, None),
    "hydro_sd_3d_rk3": (SINE, ["time/integrator=rk3"], 3, ["analytic", "mass_strict"], "W_cv_N8p3_1_0.dat", 0.1,
                        3.0e-5, None, None),
    "hydro_fv_blast_2d_rk3": (SINE, ["time/integrator=rk3", "job/fallback=true", "mesh/nx3=1",
                                     "problem/problem=spherical_blast", "hydro/gamma=1.6666666666667",
                                     "time/tlim=0.02", "output/dt=0.02"], 2, ["mass_strict"], "W_cv_N8p3_1_0.dat",
                              0.02, None, None, None),
    "hydro_implosion_2d": ("implosion.athinput", ["time/tlim=0.1", "output/dt=0.05"], 2, ["mass_strict"],
                           "W_cv_N32p3_1_0.dat", 0.1, None, None, None),
    "induction_fv_3d": ("induction_loop.athinput", [], 3, ["golden"], "B2_cv_N8p3_2_0.dat", 0.2, None,
                        "induction_fv", 1e-8),
}


@pytest.mark.parametrize("name", sorted(SUITE))
def test_reference_suite_on_gpu(tmp_path, name):
    inp, over, ndim, checks, field, t_end, l1_limit, gname, grtol = SUITE[name]
    args = ["-i", os.path.join(op.INPUTS, inp)] + over
    gpu, orc = str(tmp_path / "gpu"), str(tmp_path / "oracle")
    op.run_driver(args, gpu)
    assert os.path.isfile(os.path.join(gpu, field)), sorted(os.listdir(gpu))
    i = int(field.split("_")[-2])
    for chk in checks:
        if chk == "analytic":
            err = op.l1_vs_exact_sine(gpu, i, t_end, ndim)
            print(f"{name}: L1(rho) vs analytic = {err:.3e} (limit {l1_limit:.1e})")
            assert err < l1_limit
        elif chk == "mass_strict":
            drift = op.mass_drift(gpu)
            print(f"{name}: mass drift = {drift:.3e}")
            assert drift < 1e-12
        elif chk == "golden":
            a = np.fromfile(os.path.join(op.GOLDEN, gname, field))
            b = np.fromfile(os.path.join(gpu, field))
            assert a.shape == b.shape
            diff = np.abs(a - b).max() / max(np.abs(a).max(), 1e-300)
            print(f"{name}: golden max rel diff = {diff:.3e} (rtol {grtol:.1e}), {int((a != b).sum())} values differ")
            assert diff < grtol
    op.run_oracle(args, orc)
    res = op.compare_dirs(orc, gpu)
    bad = {k: v for k, v in res.items() if v[0]}
    assert bad == {}, bad
