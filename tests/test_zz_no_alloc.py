"""The allocation-free whole-step entry (spdk_hydro_step_fused_noalloc,
workspace in a static per-GPU array) against the allocating one
(spdk_hydro_step_fused) through the C ABI: two calls in a row, the second
on a smaller mesh, every output bit for bit; a mesh whose workspace exceeds
the array is rejected before any launch."""
import ctypes

import numpy as np
import pytest

import oracle_py as op

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import spd_k_amd as sk  # noqa: E402
from test_gpu_stages import ctx_for, dev, host, sd_shape, state  # noqa: E402


def _one(p, act, N, seed, dt):
    c, _, _, _ = ctx_for(p, act, N=N)
    n = p + 1
    s_u = sd_shape(1, 6, N, n, act)
    U0 = state(s_u, np.random.default_rng(seed))
    Ua, Ub = dev(U0), dev(U0)
    Wa = torch.zeros(s_u, dtype=torch.float64, device="cuda")
    Wb = torch.full(s_u, float("nan"), dtype=torch.float64, device="cuda")
    dta = c.hydro_step_fused(Ua, dt, Wa)
    dtb = c.hydro_step_fused_noalloc(Ub, dt, Wb)
    a, b = host(Ua), host(Ub)
    assert np.array_equal(a, b), f"U_sp: {int((a != b).sum())} of {a.size} differ"
    wa, wb = host(Wa), host(Wb)
    assert np.array_equal(wa, wb), f"W_cv: {int((wa != wb).sum())} of {wa.size} differ"
    assert dta == dtb and np.isfinite(dta) and dta > 0
    assert not np.array_equal(a, U0)  # the step changed the state
    c.close()


def test_noalloc_step_equals_allocating_step():
    cap = sk.lib().spdk_fused_noalloc_max_bytes()
    assert cap == 64 << 20
    # 3-D p=3 on 8^3 elements (10^3 with ghosts: 43.5 MB of workspace), then
    # a smaller 2-D p=2 mesh: the array keeps the first call's contents
    _one(3, (1, 1, 1), (8, 8, 8), 11, 1e-3)
    _one(2, (1, 1, 0), (6, 5, 1), 12, 2e-3)


def test_noalloc_rejects_meshes_beyond_the_array():
    c, _, _, _ = ctx_for(3, (1, 1, 1), N=(12, 12, 12))
    s = sd_shape(1, 6, (12, 12, 12), 4, (1, 1, 1))
    N = (ctypes.c_int * 3)(s[4], s[3], s[2])
    assert sk.lib().spdk_fused_workspace_bytes(c.ctx, N) > (64 << 20)
    U = torch.zeros(s, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="static scratch"):
        c.hydro_step_fused_noalloc(U, 1e-3)
    c.close()
