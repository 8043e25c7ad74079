"""The MLP oracle (CPU) against what the mathematics fixes: its exact-mode
gradient against central finite differences of the loss, and its BF16
rounding against torch's bfloat16 conversion (round to nearest even)."""
import numpy as np
import pytest
import torch

import oracle_py as op

DIMS = [7, 6, 5, 4]


def loss_of(params, X, T):
    Y = op.orc_mlp_forward(DIMS, params, X, bf16=False)
    return float(((Y - T) ** 2).mean())


def test_exact_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((9, DIMS[0]))
    T = rng.standard_normal((9, DIMS[3]))
    p = op.mlp_init(DIMS, seed=2)
    g, loss = op.orc_mlp_grad(DIMS, p, X, T, bf16=False)
    assert abs(loss - loss_of(p, X, T)) < 1e-12 * max(1.0, loss)
    # parameters are FP32: perturb with a step the FP32 value represents exactly
    for i in rng.choice(len(p), 40, replace=False):
        h = np.float32(2.0 ** -10) * max(np.float32(1.0), abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (loss_of(pp, X, T) - loss_of(pm, X, T)) / (float(pp[i]) - float(pm[i]))
        assert abs(fd - g[i]) <= 2e-5 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_gradient_of_a_perfect_fit_is_zero():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((5, DIMS[0]))
    p = op.mlp_init(DIMS, seed=4)
    T = op.orc_mlp_forward(DIMS, p, X, bf16=False)
    g, loss = op.orc_mlp_grad(DIMS, p, X, T, bf16=False)
    assert loss == 0.0 and np.abs(g).max() == 0.0


@pytest.mark.parametrize("seed", [0, 1])
def test_bf16_rounding_matches_torch(seed):
    """The oracle's BF16 rounding (FP32 then round to nearest even) against
    torch's float32 -> bfloat16 conversion over six decades."""
    import ctypes
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(2000) * 10.0 ** rng.integers(-3, 3, 2000)
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    out = np.array([op.orc().orc_bf16_round(ctypes.c_double(v)) for v in x])
    assert np.array_equal(out, ref)
