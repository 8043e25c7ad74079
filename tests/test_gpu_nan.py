"""A state that is not a number stops the run with an error instead of
running on: the dt minimum (src/hydro.cpp:208-262) and the residual and
diagnostic maxima keep a NaN, so the host's positivity / finiteness check
of the next dt fires. Finite inputs give the same bits as before (the
parity tests cover that)."""
import math

import numpy as np
import pytest

import oracle_py as op

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import spd_k_amd as sk  # noqa: E402
from test_gpu_stages import ctx_for, dev, sd_shape, state  # noqa: E402


@pytest.mark.parametrize("where", ["first", "middle", "last"])
def test_min_dt_keeps_a_planted_nan(where):
    c, m, _, _ = ctx_for(3, (1, 1, 0), N=(5, 4, 1))
    shp = sd_shape(1, 6, (5, 4, 1), 4, (1, 1, 0))
    W = state(shp, np.random.default_rng(3))
    finite = c.compute_dt(dev(W))
    assert math.isfinite(finite) and finite > 0
    # one solution point of one active element (ghost layer 1 excluded)
    k, j, i = 0, {"first": 1, "middle": 2, "last": 4}[where], {"first": 1, "middle": 3, "last": 5}[where]
    W[0, 4, k, j, i, 0, 2, 1] = np.nan
    assert math.isnan(c.compute_dt(dev(W)))
    c.close()


@pytest.mark.parametrize("extra", [
    ["job/path=stages"],
    ["job/path=fused"],
    ["job/precision=mixed"],
    ["job/precision=mixed", "job/picard=residual"],
    ["job/fallback=true"],
])
def test_run_with_nan_state_stops_with_an_error(tmp_path, extra):
    # Sod with a NaN right-hand pressure: half of the domain is not a number
    args = ["-i", f"{op.INPUTS}/sod.athinput", "mesh/nx1=16", "mesh/nx2=4", "mesh/x2_bc=periodic",
            "problem/p1=nan", "time/tlim=0.05", "output/dt=-1", "job/fallback=false"] + extra
    with pytest.raises(RuntimeError, match="physical range"):
        sk.run(args, str(tmp_path / "o"))
