"""Stage-level parity on the GPU: each kernel-layer entry point of the C ABI
(called through spd_k_amd.Context on device tensors) against the oracle's
restatement of the same reference function on the same seeded inputs.
Integer/flag outputs and FP64 values are compared bit for bit."""
import ctypes

import numpy as np
import pytest

import oracle_py as op

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import spd_k_amd as sk  # noqa: E402


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def ctx_for(p, act, bc=(0, 0, 0), gamma=1.4, N=(4, 4, 4), **kw):
    m = op.matrices(p)
    ops = {k: m[k] for k in sk.Context.MATS}
    ops["wt"] = m["wt"]
    faces, centers = [], []
    for d in range(3):
        if act[d]:
            f, c = op.fv_geometry(N[d], p)
        else:
            f, c = np.array([-1.0, 0.0]), np.array([-1.0])
        faces.append(f)
        centers.append(c)
    ops["fv_faces"], ops["fv_centers"] = faces, centers
    h = tuple(1.0 / N[d] if act[d] else 1.0 for d in range(3))
    return sk.Context(p, act, bc, p + 1, ops, gamma=gamma, h=h, **kw), m, faces, centers


def sd_shape(nt, nv, N, n, act, stag=None):
    Ne = [N[d] + 2 if act[d] else 1 for d in range(3)]
    npt = [n if act[d] else 1 for d in range(3)]
    if stag is not None and act[stag]:
        npt[stag] = n + 1
    return (nt, nv, Ne[2], Ne[1], Ne[0], npt[2], npt[1], npt[0])


def state(shape, rng, positive=True):
    """Random but physical conservative states (rho, m, E) in layout `shape`."""
    a = rng.uniform(-0.5, 0.5, size=shape)
    if positive:
        a[:, 0] = rng.uniform(0.5, 2.0, size=a[:, 0].shape)
        ke = 0.5 * (a[:, 1] ** 2 + a[:, 2] ** 2 + a[:, 3] ** 2) / a[:, 0]
        a[:, 4] = ke + rng.uniform(1.0, 3.0, size=a[:, 4].shape)
        if a.shape[1] > 5:
            a[:, 5] = 0.0
    return np.ascontiguousarray(a)


GEOMS = [((1, 1, 1), (3, 4, 2)), ((1, 1, 0), (5, 3, 1)), ((1, 0, 0), (6, 1, 1))]


@pytest.mark.parametrize("act,N", GEOMS)
@pytest.mark.parametrize("p", [1, 3, 4])
def test_transform_and_flux_riemann(act, N, p):
    rng = np.random.default_rng(10 * p + sum(act))
    c, m, _, _ = ctx_for(p, act, N=N)
    n = p + 1
    T = p + 1
    s_sp = sd_shape(T, 6, N, n, act)
    U = state(s_sp, rng)
    for d in range(3):
        if not act[d]:
            continue
        s_fp = sd_shape(T, 6, N, n, act, stag=d)
        # transform_a_to_b_1d sp -> fp
        Ufp_o = np.zeros(s_fp)
        op.orc().orc_transform_1d(op.dp(U), op.ip(s_sp), op.dp(Ufp_o), op.ip(s_fp), op.dp(m["sp_to_fp"]), n + 1, n, d, -1)
        Ug, Ufp_g = dev(U), torch.zeros(s_fp, dtype=torch.float64, device="cuda")
        c.transform_a_to_b_1d(Ug, Ufp_g, sk.SPDK_M_SP_TO_FP, d)
        assert np.array_equal(host(Ufp_g), Ufp_o)
        # compute_fluxes
        cfg = op.OrcCfg.make(act=act)
        F_o = np.zeros(s_fp)
        op.orc().orc_fluxes(ctypes.byref(cfg), op.dp(Ufp_o), op.ip(s_fp), op.dp(F_o), d)
        F_g = torch.zeros(s_fp, dtype=torch.float64, device="cuda")
        c.compute_fluxes(Ufp_g, F_g, d)
        assert np.array_equal(host(F_g), F_o)
        # boundaries (all three types) then LLF in place
        for bc in (0, 1, 2):
            cfgb = op.OrcCfg.make(act=act, bc=(bc, bc, bc))
            cb, _, _, _ = ctx_for(p, act, bc=(bc, bc, bc), N=N)
            Ub_o = Ufp_o.copy()
            op.orc().orc_sd_bc(ctypes.byref(cfgb), op.dp(Ub_o), op.ip(s_fp), d)
            Ub_g = dev(Ufp_o)
            cb.boundaries(Ub_g, d)
            assert np.array_equal(host(Ub_g), Ub_o), f"bc {bc}"
            Fr_o = F_o.copy()
            op.orc().orc_riemann(ctypes.byref(cfgb), op.dp(Ub_o), op.dp(Fr_o), op.ip(s_fp), d)
            Fr_g = dev(F_o)
            cb.sd_riemann_solver(Ub_g, Fr_g, d)
            assert np.array_equal(host(Fr_g), Fr_o), f"riemann bc {bc}"
            cb.close()
    c.close()


@pytest.mark.parametrize("act,N", GEOMS)
@pytest.mark.parametrize("p", [1, 3, 4, 5, 6])
@pytest.mark.parametrize("mat", ["sp_to_cv", "cv_to_sp"])
@pytest.mark.parametrize("nt", [1, 3])
def test_transform_a_to_b(act, N, p, mat, nt):
    """transform_a_to_b (x, y, z sweeps over the active axes) against the
    oracle's one-axis transforms applied in that order, bit for bit; 3-D p=6
    (343 points per element) takes the ping-pong sweeps, the rest the
    one-pass element kernel."""
    rng = np.random.default_rng(2000 + 10 * p + nt + sum(act))
    c, m, _, _ = ctx_for(p, act, N=N)
    n = p + 1
    shape = sd_shape(nt, 6, N, n, act)
    A = np.ascontiguousarray(rng.normal(size=shape))
    M = m[mat]
    cur = A
    for d in range(3):
        if not act[d]:
            continue
        nxt = np.zeros(shape)
        op.orc().orc_transform_1d(op.dp(cur), op.ip(shape), op.dp(nxt), op.ip(shape), op.dp(M), n, n, d, -1)
        cur = nxt
    B = torch.full(shape, float("nan"), dtype=torch.float64, device="cuda")
    T = torch.zeros(shape, dtype=torch.float64, device="cuda")
    c.transform_a_to_b(dev(A), B, T, getattr(sk, "SPDK_M_" + mat.upper()))
    got = host(B)
    assert np.array_equal(got, cur), f"{int((got != cur).sum())} of {got.size} differ"
    c.close()


@pytest.mark.parametrize("act,N", GEOMS)
@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("bc", [0, 1, 2])
def test_solve_fluxes_fused(act, N, p, bc):
    """The fused flux evaluation of given ADER slices against the oracle's
    Solve_fluxes sequence (transform sp->fp, fluxes, trace fill, LLF) per
    axis, every slice, element and point, bit for bit."""
    rng = np.random.default_rng(1000 + 10 * p + bc + sum(act))
    c, m, _, _ = ctx_for(p, act, bc=(bc, bc, bc), N=N)
    n, T = p + 1, p + 1
    s_a = sd_shape(T, 6, N, n, act)
    Ua = state(s_a, rng)
    cfg = op.OrcCfg.make(act=act, bc=(bc, bc, bc))
    F_o, F_g = [], []
    for d in range(3):
        if not act[d]:
            F_o.append(None)
            F_g.append(None)
            continue
        s_fp = sd_shape(T, 6, N, n, act, stag=d)
        Ufp = np.zeros(s_fp)
        op.orc().orc_transform_1d(op.dp(Ua), op.ip(s_a), op.dp(Ufp), op.ip(s_fp), op.dp(m["sp_to_fp"]), n + 1, n, d, -1)
        F = np.zeros(s_fp)
        op.orc().orc_fluxes(ctypes.byref(cfg), op.dp(Ufp), op.ip(s_fp), op.dp(F), d)
        op.orc().orc_sd_bc(ctypes.byref(cfg), op.dp(Ufp), op.ip(s_fp), d)
        op.orc().orc_riemann(ctypes.byref(cfg), op.dp(Ufp), op.dp(F), op.ip(s_fp), d)
        F_o.append(F)
        F_g.append(torch.full(s_fp, float("nan"), dtype=torch.float64, device="cuda"))
    c.solve_fluxes_fused(dev(Ua), F_g[0], F_g[1], F_g[2])
    for d in range(3):
        if act[d]:
            g = host(F_g[d])
            assert np.array_equal(g, F_o[d]), f"axis {d}: {int((g != F_o[d]).sum())} of {g.size} differ"
    c.close()


@pytest.mark.parametrize("act,N", GEOMS)
@pytest.mark.parametrize("p", [2, 3])
def test_predict_correct_dt(act, N, p):
    rng = np.random.default_rng(100 + p)
    c, m, _, _ = ctx_for(p, act, N=N)
    n, T = p + 1, p + 1
    s_a = sd_shape(T, 6, N, n, act)
    s_u = sd_shape(1, 6, N, n, act)
    U = state(s_u, rng)
    Ua = state(s_a, rng)
    F = [np.ascontiguousarray(rng.normal(size=sd_shape(T, 6, N, n, act, stag=d))) if act[d]
         else np.zeros(sd_shape(T, 6, N, n, act)) for d in range(3)]
    h = np.array([1.0 / N[d] if act[d] else 1.0 for d in range(3)])
    dt = 0.0123
    cfg = op.OrcCfg.make(act=act)
    Ua_o = Ua.copy()
    shapes = [op.ip(f.shape) for f in F]
    op.orc().orc_predict(ctypes.byref(cfg), op.dp(U), op.ip(s_u), op.dp(Ua_o), op.ip(s_a), op.dp(F[0]), shapes[0],
                         op.dp(F[1]), shapes[1], op.dp(F[2]), shapes[2], op.dp(m["dfp_to_sp"]), n, n + 1,
                         op.dp(m["invader"]), op.dp(m["wt"]), op.dp(h), ctypes.c_double(dt))
    Ua_g = dev(Ua)
    Fg = [dev(f) if act[d] else None for d, f in enumerate(F)]
    c.update_prediction(dev(U), Ua_g, Fg[0], Fg[1], Fg[2], p + 2, dt)
    assert np.array_equal(host(Ua_g), Ua_o)
    U_o = U.copy()
    op.orc().orc_correct(ctypes.byref(cfg), op.dp(U_o), op.ip(s_u), op.dp(Ua), op.ip(s_a), op.dp(F[0]), shapes[0],
                         op.dp(F[1]), shapes[1], op.dp(F[2]), shapes[2], op.dp(m["dfp_to_sp"]), n, n + 1,
                         op.dp(m["wt"]), op.dp(h), ctypes.c_double(dt))
    U_g = dev(U)
    c.update_solution(U_g, dev(Ua), Fg[0], Fg[1], Fg[2], p + 2, dt)
    assert np.array_equal(host(U_g), U_o)
    # compute_dt on primitive-like data
    W = U.copy()
    W[:, 4] = rng.uniform(0.5, 2.0, size=W[:, 4].shape)
    dt_o = op.orc().orc_dt(ctypes.byref(cfg), op.dp(W), op.ip(s_u), ctypes.c_double(0.8), op.dp(h), ctypes.c_double(0.0))
    assert c.compute_dt(dev(W)) == dt_o
    c.close()


def fv_shape(nv, N, n, act, stag=None):
    f = [N[d] * n + 4 if act[d] else 1 for d in range(3)]
    if stag is not None and act[stag]:
        f[stag] += 1
    return (nv, f[2], f[1], f[0])


@pytest.mark.parametrize("act,N", GEOMS)
@pytest.mark.parametrize("opts", [dict(), dict(nad_delta=1, nad_moore=0), dict(nad_delta=1), dict(sed=0, blending=0)])
def test_fv_fallback_stages(act, N, opts):
    p = 3
    n = p + 1
    rng = np.random.default_rng(7 + sum(act))
    kw = dict(nad_delta=bool(opts.get("nad_delta", 0)), nad_moore=bool(opts.get("nad_moore", 1)),
              sed=bool(opts.get("sed", 1)), blending=bool(opts.get("blending", 1)))
    c, m, faces, centers = ctx_for(p, act, N=N, **kw)
    cfg = op.OrcCfg.make(act=act, p=p, **opts)
    s_fv = fv_shape(6, N, n, act)
    # primitive fields with smooth background plus planted extrema (troubles)
    Wo = state((1,) + s_fv, rng)[0]
    Wn = Wo * (1 + 0.02 * rng.normal(size=Wo.shape))
    Wn[0].flat[rng.integers(0, Wn[0].size, 5)] *= 1.5
    Wn[4].flat[rng.integers(0, Wn[4].size, 3)] = -1e-12  # PAD
    Wo, Wn = np.ascontiguousarray(Wo), np.ascontiguousarray(Wn)
    zeros = lambda s: np.zeros(s)
    tr_o, ax_o, ay_o, az_o = zeros(s_fv), zeros(s_fv), zeros(s_fv), zeros(s_fv)
    nfc = (ctypes.c_int * 3)(*[len(x) for x in centers])
    pc = (ctypes.POINTER(ctypes.c_double) * 3)(*[op.dp(x) for x in centers])
    pf = (ctypes.POINTER(ctypes.c_double) * 3)(*[op.dp(x) for x in faces])
    op.orc().orc_detect(ctypes.byref(cfg), op.dp(Wn), op.dp(Wo), op.dp(tr_o), op.dp(ax_o), op.dp(ay_o), op.dp(az_o),
                        op.ip(s_fv), pc, pf, nfc, 1, 0b10001)
    tr_g = torch.zeros(s_fv, dtype=torch.float64, device="cuda")
    al = [torch.zeros(s_fv, dtype=torch.float64, device="cuda") for _ in range(3)]
    c.detect_troubles(dev(Wn), dev(Wo), tr_g, al[0], al[1], al[2], 1, 0b10001)
    assert np.array_equal(host(tr_g), tr_o)
    assert tr_o[5].sum() > 0
    # the same troubles with the SED coefficients as bit words, twice in a
    # row on one word array (the driver's use: words persist across calls)
    tr_b = torch.zeros(s_fv, dtype=torch.float64, device="cuda")
    abits = torch.zeros(s_fv[1:], dtype=torch.int16, device="cuda")
    for _ in range(2):
        c.detect_troubles_bits(dev(Wn), dev(Wo), tr_b, abits, 1, 0b10001)
        assert np.array_equal(host(tr_b), tr_o)
    # FV ghosts of the flags, then theta
    op.orc().orc_fv_bc(ctypes.byref(cfg), op.dp(tr_o), op.ip(s_fv))
    c.fv_boundaries(tr_g)
    assert np.array_equal(host(tr_g), tr_o)
    s_th = (1,) + s_fv[1:]
    th_tmp_o, th_o = zeros(s_th), zeros(s_th)
    op.orc().orc_blend(ctypes.byref(cfg), op.dp(tr_o), op.dp(th_tmp_o), op.dp(th_o), op.ip(s_fv))
    th_tmp_g, th_g = torch.zeros(s_th, dtype=torch.float64, device="cuda"), torch.zeros(s_th, dtype=torch.float64, device="cuda")
    c.blend_theta(tr_g, th_tmp_g, th_g)
    assert np.array_equal(host(th_g), th_o)
    op.orc().orc_fv_bc(ctypes.byref(cfg), op.dp(th_o), op.ip(s_th))
    c.fv_boundaries(th_g)
    # blended MUSCL-Hancock + HLLC face fluxes
    Fs = [np.ascontiguousarray(rng.normal(size=fv_shape(6, N, n, act, stag=d))) for d in range(3)]
    Fo = [f.copy() for f in Fs]
    op.orc().orc_fallback(ctypes.byref(cfg), op.dp(Wo), op.dp(th_o), op.dp(Fo[0]), op.dp(Fo[1]), op.dp(Fo[2]),
                          op.ip(s_fv), pc, pf, nfc, ctypes.c_double(1e-3))
    Fg = [dev(f) for f in Fs]
    c.fallback_fluxes(dev(Wo), th_g, Fg[0], Fg[1] if act[1] else None, Fg[2] if act[2] else None, 1e-3)
    for d in range(3):
        if act[d]:
            assert np.array_equal(host(Fg[d]), Fo[d]), d
    # the FV update pair of one slice: update = 0 with the SD face fluxes, then
    # the blended update (recomputing only cells next to a nonzero theta)
    # against update = 1 with the blended fluxes; theta nonzero on a few
    # cells, the fluxes changed on exactly their faces
    s_cv = sd_shape(1, 6, N, n, act)
    Ucv = state(s_cv, rng)
    th_s = np.zeros(s_th)
    lo = [2 if act[d] else 0 for d in range(3)]
    cells = [tuple(int(rng.integers(lo[d], s_fv[3 - d] - lo[d])) for d in (2, 1, 0)) for _ in range(4)]
    Fb = [f.copy() for f in Fs]
    for K, J, I in cells:
        th_s[0, K, J, I] = 0.5
        for d in range(3):
            if act[d]:
                for off in (0, 1):
                    KK, JJ, II = K + off * (d == 2), J + off * (d == 1), I + off * (d == 0)
                    Fb[d][:, KK, JJ, II] += rng.normal(size=6)
    assert (th_s == 0).sum() > th_s.size // 2
    Fsd = [dev(f) if act[d] else None for d, f in enumerate(Fs)]
    Fbl = [dev(f) if act[d] else None for d, f in enumerate(Fb)]
    th_g = dev(th_s)
    t_id, dt = 1, 0.0137
    Un_a, Uo_a, Ucv_a = (torch.zeros(s_fv, dtype=torch.float64, device="cuda"),
                         torch.zeros(s_fv, dtype=torch.float64, device="cuda"), dev(Ucv))
    c.fv_update_solution(Un_a, Uo_a, Ucv_a, Fsd[0], Fsd[1], Fsd[2], t_id, dt, 0)
    c.fv_update_blended(Un_a, Ucv_a, Fbl[0], Fbl[1], Fbl[2], th_g, t_id, dt)
    Un_b, Uo_b, Ucv_b = (torch.zeros(s_fv, dtype=torch.float64, device="cuda"),
                         torch.zeros(s_fv, dtype=torch.float64, device="cuda"), dev(Ucv))
    c.fv_update_solution(Un_b, Uo_b, Ucv_b, Fbl[0], Fbl[1], Fbl[2], t_id, dt, 1)
    assert np.array_equal(host(Un_a), host(Un_b))
    assert np.array_equal(host(Ucv_a), host(Ucv_b))
    assert np.array_equal(host(Uo_a), host(Uo_b))
    c.close()


@pytest.mark.parametrize("act,N", GEOMS)
@pytest.mark.parametrize("bc", [0, 1, 2])
def test_fv_update_prims(act, N, bc):
    """The update = 0 FV update fused with the primitives of U_new and U_old,
    ghosts filled on W, against the reference's order (update = 0, ghosts on
    U, then compute_primitives), every cell bit for bit, all boundary types."""
    p, n = 3, 4
    rng = np.random.default_rng(300 + bc + sum(act))
    c, _, _, _ = ctx_for(p, act, bc=(bc, bc, bc), N=N)
    s_fv = fv_shape(6, N, n, act)
    s_cv = sd_shape(1, 6, N, n, act)
    Ucv = dev(state(s_cv, rng))
    F = [dev(0.01 * rng.normal(size=fv_shape(6, N, n, act, stag=d))) if act[d] else None for d in range(3)]
    z = lambda: torch.zeros(s_fv, dtype=torch.float64, device="cuda")
    Un_a, Uo_a, Wn_a, Wo_a = z(), z(), z(), z()
    c.fv_update_solution(Un_a, Uo_a, Ucv, F[0], F[1], F[2], 2, 0.01, 0)
    c.fv_boundaries(Uo_a)
    c.fv_boundaries(Un_a)
    c.compute_primitives_fv(Un_a, Wn_a)
    c.compute_primitives_fv(Uo_a, Wo_a)
    Un_b, Wn_b, Wo_b = z(), z(), z()
    c.fv_update_prims(Un_b, Wn_b, Wo_b, Ucv, F[0], F[1], F[2], 2, 0.01)
    c.fv_boundaries(Wo_b)
    c.fv_boundaries(Wn_b)
    c.fv_boundaries(Un_b)
    for name, x, y in (("W_new", Wn_a, Wn_b), ("W_old", Wo_a, Wo_b), ("U_new", Un_a, Un_b)):
        a, b = host(x), host(y)
        assert np.array_equal(a, b), f"{name}: {int((a != b).sum())} of {a.size} differ"
        assert np.array_equal(np.signbit(a), np.signbit(b)), name
    c.close()



@pytest.mark.parametrize("act,N", [((1, 0, 0), (5, 1, 1)), ((1, 1, 0), (4, 3, 1)), ((1, 1, 1), (3, 3, 2))])
@pytest.mark.parametrize("p", [1, 3])
def test_viscous_stages(act, N, p):
    """The viscous stage entries against the oracle on the same inputs, bit
    for bit: sd_riemann_solver with the viscous flag (F and the replaced
    face states), compute_gradient, compute_viscous_flux (nu, beta) and
    sd_rusanov_solver, every axis (src/hydro.cpp:368-555)."""
    rng = np.random.default_rng(700 + 10 * p + sum(act))
    nu, beta = 0.003, -0.0017
    c, m, _, _ = ctx_for(p, act, N=N, nu=nu, beta=beta)
    n, T = p + 1, p + 1
    cfg = op.OrcCfg.make(act=act)
    L = op.orc()
    L.orc_gradient.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_viscous_flux.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    h = [1.0 / N[d] if act[d] else 1.0 for d in range(3)]
    s_sp = sd_shape(T, 6, N, n, act)
    dU_o = [np.zeros(s_sp) for _ in range(3)]
    dU_g = [torch.zeros(s_sp, dtype=torch.float64, device="cuda") for _ in range(3)]
    Ufp = {}
    for d in range(3):
        if not act[d]:
            continue
        s_fp = sd_shape(T, 6, N, n, act, stag=d)
        U = state(s_fp, rng)
        F = rng.normal(size=s_fp)
        # Riemann with the viscous flag
        Uo, Fo = U.copy(), F.copy()
        L.orc_riemann_viscous(ctypes.byref(cfg), op.dp(Uo), op.dp(Fo), op.ip(s_fp), d)
        Ug, Fg = dev(U), dev(F)
        c.sd_riemann_solver(Ug, Fg, d, viscous=True)
        assert np.array_equal(host(Fg), Fo), f"riemann F axis {d}"
        assert np.array_equal(host(Ug), Uo), f"riemann U axis {d}"
        # gradient of those face-updated values
        L.orc_gradient(ctypes.byref(cfg), op.dp(Uo), op.ip(s_fp), op.dp(dU_o[d]), op.ip(s_sp), h[d],
                       op.dp(m["dfp_to_sp"]), n, n + 1, d)
        c.compute_gradient(Ug, dU_g[d], d)
        assert np.array_equal(host(dU_g[d]), dU_o[d]), f"gradient axis {d}"
        Ufp[d] = (Uo, Ug, s_fp, Fo, Fg)
    for d, (Uo, Ug, s_fp, Fo, Fg) in Ufp.items():
        o = [dU_o[d], dU_o[(d + 1) % 3], dU_o[(d + 2) % 3]]
        g = [dU_g[d], dU_g[(d + 1) % 3], dU_g[(d + 2) % 3]]
        L.orc_viscous_flux(ctypes.byref(cfg), op.dp(Uo), op.ip(s_fp), op.dp(o[0]), op.dp(o[1]), op.dp(o[2]),
                           op.ip(s_sp), op.dp(m["sp_to_fp"]), n + 1, n, nu, beta, d)
        c.compute_viscous_flux(Ug, g[0], g[1], g[2], d)
        assert np.array_equal(host(Ug), Uo), f"viscous flux axis {d}"
        L.orc_rusanov(op.dp(Uo), op.dp(Fo), op.ip(s_fp), d)
        c.sd_rusanov_solver(Ug, Fg, d)
        assert np.array_equal(host(Fg), Fo), f"rusanov axis {d}"
    c.close()
