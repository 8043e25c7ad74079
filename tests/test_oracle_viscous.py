"""Pins of the oracle's restatement of the reference's viscous terms
(src/hydro.cpp:265-268, :368-555; sequencing src/hydro_ader.hpp:278-289,
:380-405) against what the mathematics and the Navier-Stokes stress fix:
the gradient is exact on polynomials, the viscous flux of a uniform velocity
gradient is the Newtonian stress (Stokes' bulk term with beta = -2/3 nu) and
its work, the Rusanov step adds the right element's face value, the viscous
Riemann call leaves F as the inviscid one and puts the right state into
both face copies, and a uniform state stays uniform with viscosity on.

The reference's operator adds the viscous flux at element faces only (the
interior flux points keep the inviscid flux) and with a + sign; run on the
diffusing shear pulse of its user_ic.hpp it does not follow the closed form
sigma^2 = sigma0^2 + 2 nu t (measured here and recorded, not asserted)."""
import ctypes
import os

import numpy as np
import pytest

import oracle_py as op

USER = os.path.join(op.INPUTS, "user.athinput")


def _ip(shape):
    return op.ip(shape)


def _setup(p, act):
    L = op.orc()
    for f in ("orc_gradient", "orc_viscous_flux", "orc_rusanov", "orc_riemann_viscous"):
        getattr(L, f).restype = ctypes.c_int
    L.orc_gradient.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_viscous_flux.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    cfg = op.OrcCfg.make(act=act, p=p)
    return L, cfg, op.matrices(p)


def _shapes(nt, N, n, act, stag):
    Ne = [N[d] + 2 if act[d] else 1 for d in range(3)]
    sp = [n if act[d] else 1 for d in range(3)]
    fp = list(sp)
    fp[stag] = n + 1
    return ((nt, 6, Ne[2], Ne[1], Ne[0], sp[2], sp[1], sp[0]), (nt, 6, Ne[2], Ne[1], Ne[0], fp[2], fp[1], fp[0]))


@pytest.mark.parametrize("p", [1, 3, 5])
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_gradient_exact_on_polynomials(p, dim):
    """compute_gradient of flux-point values of a polynomial of degree p+1
    along `dim` equals its derivative at the solution points (active
    elements; ghost elements untouched)."""
    act = (1, 1, 1)
    N = (3, 2, 2)
    L, cfg, m = _setup(p, act)
    n = p + 1
    s_sp, s_fp = _shapes(2, N, n, act, dim)
    h = 0.25
    rng = np.random.default_rng(p + 10 * dim)
    coef = rng.normal(size=(2, 6, p + 2))  # per slice and variable
    U = np.zeros(s_fp)
    want = np.zeros(s_sp)
    Ne = s_sp[2:5][::-1]  # (Nx, Ny, Nz)
    for e in np.ndindex(Ne[2], Ne[1], Ne[0]):
        ed = e[2 - dim]
        xf = (ed - 1 + m["x_fp"]) * h
        xs = (ed - 1 + m["x_sp"]) * h
        for t in range(2):
            for v in range(6):
                P = np.polynomial.Polynomial(coef[t, v])
                vals, der = P(xf), P.deriv()(xs)
                sl_f = [slice(None)] * 3
                sl_s = [slice(None)] * 3
                shape_f = [1, 1, 1]
                shape_f[2 - dim] = n + 1
                shape_s = [1, 1, 1]
                shape_s[2 - dim] = n
                U[(t, v) + e] = np.broadcast_to(vals.reshape(shape_f), s_fp[5:])
                want[(t, v) + e] = np.broadcast_to(der.reshape(shape_s), s_sp[5:])
    dU = np.full(s_sp, 7.0)
    L.orc_gradient(ctypes.byref(cfg), op.dp(U), _ip(s_fp), op.dp(dU), _ip(s_sp), h, op.dp(m["dfp_to_sp"]), n, n + 1, dim)
    inner = (slice(None), slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    scale = np.abs(want[inner]).max()
    assert np.abs(dU[inner] - want[inner]).max() <= 1e-11 * scale
    # the ghost elements keep their values
    ghost = dU.copy()
    ghost[inner] = 7.0
    assert np.all(ghost == 7.0)


def _flux_of_gradient(dim, du, w_vel, nu, beta):
    """Viscous flux the oracle computes at the flux points of a 3-D mesh for
    constant conserved-variable gradients du[d] (rho = 1, velocity w_vel)."""
    act = (1, 1, 1)
    N = (2, 2, 2)
    p = 2
    L, cfg, m = _setup(p, act)
    n = p + 1
    s_sp, s_fp = _shapes(1, N, n, act, dim)
    U = np.zeros(s_fp)
    U[:, 0] = 1.0
    for a in range(3):
        U[:, 1 + a] = w_vel[a]
    U[:, 4] = 2.5 + 0.5 * sum(x * x for x in w_vel)
    G = [np.zeros(s_sp) for _ in range(3)]
    for d in range(3):  # gradient along axis d of every conserved variable
        for v in range(6):
            G[d][:, v] = du[d][v]
    order = [G[dim], G[(dim + 1) % 3], G[(dim + 2) % 3]]
    L.orc_viscous_flux(ctypes.byref(cfg), op.dp(U), _ip(s_fp), op.dp(order[0]), op.dp(order[1]), op.dp(order[2]),
                       _ip(s_sp), op.dp(m["sp_to_fp"]), n + 1, n, nu, beta, dim)
    return U


@pytest.mark.parametrize("dim", [0, 1, 2])
def test_viscous_flux_is_the_newtonian_stress(dim):
    """rho = 1 and a constant velocity gradient g_ab = d v_a / d x_b: the flux
    along `dim` is the Newtonian stress row tau_{a,dim} = nu (g_{a,dim} +
    g_{dim,a}) + beta delta_{a,dim} div v (beta = -2/3 nu, Stokes) in the
    momentum slots, the stress work v . tau in the energy slot and zero mass
    flux, at every flux point."""
    nu = 0.01
    beta = -2.0 / 3.0 * nu
    rng = np.random.default_rng(3 + dim)
    g = rng.normal(size=(3, 3))  # g[a][b] = d v_a / d x_b
    vel = rng.normal(size=3)
    # conserved gradients with rho = 1: d(rho v_a)/dx_b = g_ab, d rho = 0, energy gradient unused
    du = [[0.0, g[0][b], g[1][b], g[2][b], 0.3, 0.0] for b in range(3)]
    U = _flux_of_gradient(dim, du, vel, nu, beta)
    div = g[0][0] + g[1][1] + g[2][2]
    tau = np.array([nu * (g[a][dim] + g[dim][a]) + (beta * div if a == dim else 0.0) for a in range(3)])
    want = np.array([0.0, tau[0], tau[1], tau[2], float(vel @ tau), 0.0])
    got = np.moveaxis(U[0], 0, -1).reshape(-1, 6)
    assert np.allclose(got, want[None, :], rtol=1e-13, atol=1e-15), (got[0], want)


def test_rusanov_adds_the_right_face_value():
    """sd_rusanov_solver: at every element face both copies of F grow by the
    right element's face value (riemann_wind(-0.5) of the right and left
    values, which is the right value up to rounding), interior points
    untouched."""
    L, cfg, m = _setup(2, (1, 1, 0))
    s_sp, s_fp = _shapes(2, (3, 2, 1), 3, (1, 1, 0), 0)
    rng = np.random.default_rng(5)
    U = rng.normal(size=s_fp)
    F0 = rng.normal(size=s_fp)
    F = F0.copy()
    L.orc_rusanov(op.dp(U), op.dp(F), _ip(s_fp), 0)
    for i in range(s_fp[4] - 1):
        right, left = U[..., i + 1, :, :, 0], U[..., i, :, :, -1]
        f = 0.5 * (left + right) + (-0.5) * (left - right)
        assert np.allclose(f, right, rtol=0, atol=4e-16 * np.abs(U).max())
        assert np.array_equal(F[..., i + 1, :, :, 0] - F0[..., i + 1, :, :, 0] + F0[..., i + 1, :, :, 0],
                              F0[..., i + 1, :, :, 0] + f)
        assert np.array_equal(F[..., i, :, :, -1], F0[..., i, :, :, -1] + f)
    assert np.array_equal(F[..., 1:-1], F0[..., 1:-1])


def test_riemann_viscous_keeps_F_and_sets_face_states():
    """sd_riemann_solver(viscous): F equals the inviscid call's bit for bit
    and both copies of each interior face state become riemann_wind(0.5) of
    the two, i.e. the right element's state."""
    L, cfg, m = _setup(2, (1, 0, 0))
    s_sp, s_fp = _shapes(1, (4, 1, 1), 3, (1, 0, 0), 0)
    rng = np.random.default_rng(6)
    U = rng.uniform(0.5, 1.5, size=s_fp)
    U[:, 4] += 3.0
    U[:, 5] = 0.0
    F1, F2 = np.zeros(s_fp), np.zeros(s_fp)
    U2 = U.copy()
    L.orc_riemann(ctypes.byref(cfg), op.dp(U), op.dp(F1), _ip(s_fp), 0)
    L.orc_riemann_viscous(ctypes.byref(cfg), op.dp(U2), op.dp(F2), _ip(s_fp), 0)
    assert np.array_equal(F1, F2)
    for i in range(s_fp[4] - 1):
        a, b = U[..., i, :, :, -1], U[..., i + 1, :, :, 0]
        assert np.allclose(U2[..., i, :, :, -1], b, rtol=1e-15, atol=1e-15)
        assert np.array_equal(U2[..., i, :, :, -1], U2[..., i + 1, :, :, 0])
        assert np.array_equal(U2[..., i, :, :, -1], 0.5 * (b + a) + 0.5 * (b - a))


def test_uniform_state_stays_uniform_with_viscosity(tmp_path):
    """A uniform flow with viscosity on (ADER, 2-D): every gradient is zero
    up to the rounding of the derivative rows' sums, so the viscous fluxes
    vanish and the state stays uniform to round-off."""
    args = ["-i", USER, "mesh/nx1=6", "mesh/nx2=5", "job/fallback=false", "hydro/nu=0.01", "problem/amp=0",
            "time/tlim=0.02", "output/dt=0.01"]
    op.run_oracle(args, str(tmp_path))
    g = op.grid(str(tmp_path))
    i = op.output_indices(str(tmp_path))[-1]
    for v in range(5):
        x = op.load_sd(str(tmp_path), g, i, v)
        x0 = op.load_sd(str(tmp_path), g, 0, v)
        assert np.abs(x - x0).max() <= 1e-13 * max(np.abs(x0).max(), 1.0), v


def test_shear_pulse_measured(tmp_path):
    """The reference's verification case (src/user_ic.hpp:19-26): a Gaussian
    shear pulse in v_y, nu = 5e-3, 1-D, 32 elements, p=3, a few steps. The
    restated operator stays finite and conserves mass to round-off; its
    second moment is recorded next to the closed form
    sigma^2 = sigma0^2 + 2 nu t (printed; the reference's face-only, added
    viscous flux does not follow it)."""
    args = ["-i", USER, "mesh/nx1=32", "mesh/nx2=1", "mesh/nx3=1", "job/fallback=false", "hydro/nu=0.005",
            "problem/amp=0.5", "problem/sigma=0.1", "problem/radius=0.5", "problem/v1=1.0", "time/tlim=0.005",
            "output/dt=0.005"]
    d = str(tmp_path)
    op.run_oracle(args, d)
    g = op.grid(d)
    xf = g["faces"][0]
    xc, w = 0.5 * (xf[1:] + xf[:-1]), np.diff(xf)
    out = []
    for i in op.output_indices(d):
        vy = op.load_sd(d, g, i, 2)[0, 0, :]
        m = float((vy * w).sum())
        xm = float((xc * vy * w).sum() / (vy * w).sum())
        out.append((m, float(((xc - xm) ** 2 * vy * w).sum() / (vy * w).sum()), float(vy.max())))
        assert np.isfinite(vy).all()
    (m0, s0, _), (m1, s1, pk) = out[0], out[-1]
    print(f"sigma^2: {s0:.6f} -> {s1:.6f} (closed form {s0 + 2 * 0.005 * 0.005:.6f}); peak {pk:.4f}")
    assert op.mass_drift(d) < 1e-12
