"""The MLP kernels (BF16 MFMA, FP32 accumulation) against the CPU oracle in
its BF16 mode: forward outputs and all parameter gradients, on dimensions
and a batch that are not multiples of the 64-wide tiles; and GPU training
(forward, backward and Adam on the device) fitting a teacher network."""
import ctypes

import numpy as np
import pytest
import torch

import oracle_py as op
import spd_k_amd

pytestmark = pytest.mark.gpu


def make_net(dims, cap, params):
    L = spd_k_amd.lib()
    net = ctypes.c_void_p()
    spd_k_amd.check(L.spdk_mlp_create((ctypes.c_int * 4)(*dims), cap, None, ctypes.byref(net)), "spdk_mlp_create")
    spd_k_amd.check(L.spdk_mlp_set_params(net, params.ctypes.data_as(ctypes.c_void_p)), "set_params")
    return net


def dev(a):
    return torch.tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda:0")


DIMS = [70, 128, 64, 90]


def test_forward_matches_oracle():
    rng = np.random.default_rng(5)
    B = 100
    p = op.mlp_init(DIMS, seed=6)
    X = rng.standard_normal((B, DIMS[0]))
    net = make_net(DIMS, 128, p)
    L = spd_k_amd.lib()
    Xd, Yd = dev(X), torch.zeros((B, DIMS[3]), device="cuda:0")
    spd_k_amd.check(L.spdk_mlp_forward(net, ctypes.c_void_p(Xd.data_ptr()), ctypes.c_long(DIMS[0]), B,
                                       ctypes.c_void_p(Yd.data_ptr()), ctypes.c_long(DIMS[3])), "forward")
    torch.cuda.synchronize()
    Y = Yd.cpu().double().numpy()
    ref = op.orc_mlp_forward(DIMS, p, X, bf16=True)
    exact = op.orc_mlp_forward(DIMS, p, X, bf16=False)
    # BF16 rounding of the hidden activations can differ by one unit where the
    # FP32 and FP64 pre-activations straddle a rounding boundary
    assert np.abs(Y - ref).max() <= 1e-2 * np.abs(ref).max()
    assert np.linalg.norm(Y - ref) <= 2e-3 * np.linalg.norm(ref)
    # and the BF16 forward is close to the exact one
    assert np.linalg.norm(Y - exact) <= 3e-2 * np.linalg.norm(exact)
    L.spdk_mlp_destroy(net)


def test_gradients_match_oracle():
    rng = np.random.default_rng(7)
    B = 100
    p = op.mlp_init(DIMS, seed=8)
    X = rng.standard_normal((B, DIMS[0]))
    T = rng.standard_normal((B, DIMS[3]))
    net = make_net(DIMS, 128, p)
    L = spd_k_amd.lib()
    n = op.mlp_param_count(DIMS)
    g = np.zeros(n, dtype=np.float32)
    loss = ctypes.c_float()
    Xd, Td = dev(X), dev(T)
    spd_k_amd.check(L.spdk_mlp_gradients(net, ctypes.c_void_p(Xd.data_ptr()), ctypes.c_long(DIMS[0]),
                                         ctypes.c_void_p(Td.data_ptr()), ctypes.c_long(DIMS[3]), B,
                                         g.ctypes.data_as(ctypes.c_void_p), ctypes.byref(loss)), "gradients")
    ref, ref_loss = op.orc_mlp_grad(DIMS, p, X, T, bf16=True)
    assert abs(loss.value - ref_loss) <= 1e-3 * ref_loss
    o = 0
    for l in range(3):
        for size in (DIMS[l + 1] * DIMS[l], DIMS[l + 1]):
            a, b = g[o:o + size].astype(np.float64), ref[o:o + size]
            assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b), (l, size, np.linalg.norm(a - b) / np.linalg.norm(b))
            o += size
    L.spdk_mlp_destroy(net)


def test_training_fits_a_teacher():
    """Adam on the device drives the loss on a teacher network's outputs down
    by more than 20x from a different random initialisation."""
    rng = np.random.default_rng(9)
    dims = [32, 128, 128, 48]
    B = 512
    teacher = op.mlp_init(dims, seed=10, scale=0.5)
    X = rng.standard_normal((B, dims[0]))
    T = op.orc_mlp_forward(dims, teacher, X, bf16=False)
    net = make_net(dims, B, op.mlp_init(dims, seed=11, scale=0.5))
    L = spd_k_amd.lib()
    Xd, Td = dev(X), dev(T)
    loss = ctypes.c_float()
    losses = []
    for it in range(600):
        spd_k_amd.check(L.spdk_mlp_train_step(net, ctypes.c_void_p(Xd.data_ptr()), ctypes.c_long(dims[0]),
                                              ctypes.c_void_p(Td.data_ptr()), ctypes.c_long(dims[3]), B,
                                              ctypes.c_float(2e-3), ctypes.byref(loss) if it % 50 == 0 else None),
                        "train_step")
        if it % 50 == 0:
            losses.append(loss.value)
    print(losses)
    assert losses[-1] < losses[0] / 20, losses
    L.spdk_mlp_destroy(net)
