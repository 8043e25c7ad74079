"""Pins of the oracle's FV-fallback chain (trouble detection, theta blending,
MUSCL-Hancock + HLLC fallback fluxes) and of its boundary rules (SD face
traces, FV ghost layers), against expectations derived by hand from the
mathematics and from the reference's documented index rules, never from the
oracle's own code:

- NAD (src/trouble_detection.cpp:35-104): fields with W_new = W_old are never
  flagged; a planted spike is flagged at exactly its cell; a diagonal-only
  extremum is inside the Moore box but outside the face neighbourhood; the
  relative and delta bands differ where the hand-computed bands differ;
  the loop covers FV cells 1..N-2.
- SED (src/trouble_detection.cpp:107-148): on a uniform grid a quadratic has
  d2u = 2 and both one-sided derivative jumps 2h against dv = h, so alpha = 1
  at every cell where it is computed (2..N-3); a step keeps alpha < 1.
- relax (src/trouble_detection.cpp:152-183): flags of a smooth field are
  cleared, except within one cell of the never-written alpha at 0, 1, N-2,
  N-1 (the first and last two active rows keep their flags).
- PAD (src/trouble_detection.cpp:244-259): bounds 1e-10 and 1e10 inclusive.
- theta (src/trouble_detection.cpp:191-242): weights 1, .75, .5, .375 by the
  number of nonzero offsets, then a 0.25 ring one box further out.
- fallback (src/hydro.cpp:591-788, src/muscl.hpp:2-13): theta = 0 leaves the
  SD flux untouched; faces 2..N-2 of every axis are the written set; a uniform
  state gives the physical Euler flux; a linear profile is reconstructed
  exactly at the faces; the Hancock step moves a linear density by
  -dt/2 (v . grad rho), transverse term included; the face weight is the max
  of the two adjacent cells' theta.
- boundaries (src/boundary.cpp:48-190): the reflective wall's Riemann problem
  carries zero mass, transverse-momentum and energy flux; the gradfree wall's
  flux is the physical flux of the interior trace; periodic walls share one
  face flux bit for bit; FV ghost layers on a literal 1-D example.
- whole runs: MUSCL-Hancock + HLLC alone (theta = 1 everywhere, forced with a
  negative NAD tolerance) converges to the exact Sod solution; a blast
  symmetric about the centre gives the same result with reflective and with
  periodic walls (the walls are symmetry planes of the periodic solution).
"""
import ctypes
import math
import os

import numpy as np
import pytest

import oracle_py as op

GM = 1.4
RHO, M1, M2, M3, EN, SLOT = range(6)
P = 4  # pressure index of a primitive state
MASK = (1 << RHO) | (1 << P)  # hydro limits density and pressure (src/hydro_ader.hpp:524)


# ---------------------------------------------------------------- helpers
def _ptrs(arrays):
    return (ctypes.POINTER(ctypes.c_double) * 3)(*[op.dp(a) for a in arrays])


def _uniform_axis(n, L=1.0):
    f = np.linspace(0.0, L, n + 1)
    return f, 0.5 * (f[1:] + f[:-1])


def _geom(shape, act, real_p=None, N=None):
    """Per-axis FV centres/faces (x, y, z) for an FV array of `shape` (nv, Nz, Ny, Nx):
    uniform unit spacing, or the solver's sub-grid of N elements of degree real_p."""
    cells = (shape[3], shape[2], shape[1])
    C, F = [], []
    for d in range(3):
        if act[d] and real_p is not None:
            f, c = op.fv_geometry(N[d], real_p)
        elif act[d]:
            f, c = _uniform_axis(cells[d], float(cells[d]))
        else:
            f, c = np.array([0.0, 1.0]), np.array([0.5])
        assert len(c) == cells[d]
        C.append(np.ascontiguousarray(c, dtype=np.float64))
        F.append(np.ascontiguousarray(f, dtype=np.float64))
    return C, F


def detect(Wn, Wo, act, pad=0, mask=MASK, sentinel=7.0, geom=None, **cfg):
    cfg.setdefault("sed", 0)
    c = op.OrcCfg.make(act=act, **cfg)
    tr = np.full(Wn.shape, sentinel)
    al = [np.zeros(Wn.shape) for _ in range(3)]
    C, F = geom if geom is not None else _geom(Wn.shape, act)
    nfc = (ctypes.c_int * 3)(*[len(x) for x in C])
    op.orc().orc_detect(ctypes.byref(c), op.dp(np.ascontiguousarray(Wn)), op.dp(np.ascontiguousarray(Wo)), op.dp(tr),
                        op.dp(al[0]), op.dp(al[1]), op.dp(al[2]), op.ip(Wn.shape), _ptrs(C), _ptrs(F), nfc, pad, mask)
    return tr, al


def blend(tr, act, blending=1, sentinel=0.0):
    """theta_tmp starts at zero as the solver allocates it; theta at `sentinel`."""
    c = op.OrcCfg.make(act=act, blending=blending)
    s1 = (1,) + tr.shape[1:]
    tmp, th = np.zeros(s1), np.full(s1, sentinel)
    op.orc().orc_blend(ctypes.byref(c), op.dp(np.ascontiguousarray(tr)), op.dp(tmp), op.dp(th), op.ip(tr.shape))
    return tmp[0], th[0]


def fallback(W, th, act, dt, F0=None, geom=None):
    c = op.OrcCfg.make(act=act, gamma=GM)
    nv, Nz, Ny, Nx = W.shape
    shapes = [(nv, Nz, Ny, Nx + act[0]), (nv, Nz, Ny + act[1], Nx), (nv, Nz + act[2], Ny, Nx)]
    F = [np.zeros(s) if F0 is None else np.ascontiguousarray(F0[d]).copy() for d, s in enumerate(shapes)]
    C, Fc = geom if geom is not None else _geom(W.shape, act)
    nfc = (ctypes.c_int * 3)(*[len(x) for x in C])
    op.orc().orc_fallback(ctypes.byref(c), op.dp(np.ascontiguousarray(W)),
                          op.dp(np.ascontiguousarray(th, dtype=np.float64).reshape((1,) + W.shape[1:])),
                          op.dp(F[0]), op.dp(F[1]), op.dp(F[2]), op.ip(W.shape), _ptrs(C), _ptrs(Fc), nfc,
                          ctypes.c_double(dt))
    return F


def fv_shape(act, n_act=8):
    """(nv, Nz, Ny, Nx) with n_act cells per active axis plus 2 ghost layers each side."""
    return (6,) + tuple(n_act + 4 if act[d] else 1 for d in (2, 1, 0))


def interior(shape, act, lo=1):
    """Slices of FV cells lo..N-1-lo along every active axis (all cells along inactive ones)."""
    return tuple(slice(lo, shape[1 + a] - lo) if act[2 - a] else slice(None) for a in range(3))


def prim_state(shape, rho=1.0, v=(0.0, 0.0, 0.0), p=1.0):
    W = np.zeros(shape)
    W[RHO], W[P] = rho, p
    for d in range(3):
        W[1 + d] = v[d]
    return W


def euler_flux(w, d):
    """Physical Euler flux along axis d of one primitive state (textbook form)."""
    rho, v, p = w[0], np.array(w[1:4]), w[4]
    E = p / (GM - 1.0) + 0.5 * rho * float(v @ v)
    f = np.zeros(6)
    f[0] = rho * v[d]
    f[1:4] = rho * v[d] * v
    f[1 + d] += p
    f[4] = (E + p) * v[d]
    return f


# ---------------------------------------------------------------- NAD
ACTS = [(1, 0, 0), (1, 1, 0), (1, 1, 1)]


@pytest.mark.parametrize("act", ACTS)
@pytest.mark.parametrize("moore,delta", [(1, 0), (0, 0), (1, 1), (0, 1)])
def test_nad_unchanged_fields_are_never_flagged(act, moore, delta):
    """u_new = u_old lies inside [min, max] of any neighbourhood containing
    the cell itself: no flag for constant, linear or rough fields."""
    shp = fv_shape(act, 6)
    rng = np.random.default_rng(sum(act) + 10 * moore + 100 * delta)
    for W in (prim_state(shp, 2.0, (0.3, 0, 0), 3.0), np.abs(rng.normal(1.0, 0.5, shp)) + 0.1):
        if W is not None and W.ndim == 4 and act == (1, 0, 0):
            W[RHO] += np.arange(shp[3])[None, None, :] * 0.1  # a linear part
        tr, _ = detect(W, W, act, nad_moore=moore, nad_delta=delta)
        ix = (slice(None),) + interior(shp, act)
        assert np.all(tr[(RHO,) + ix[1:]] == 0) and np.all(tr[(P,) + ix[1:]] == 0)
        assert np.all(tr[(SLOT,) + ix[1:]] == 0)


@pytest.mark.parametrize("act", ACTS)
@pytest.mark.parametrize("var", [RHO, P, M1])
def test_nad_flags_exactly_a_planted_spike(act, var):
    """One cell of W_new raised by 10%: flagged there only, and only when the
    variable is limited (density, pressure); the aggregate slot follows."""
    shp = fv_shape(act, 6)
    Wo = prim_state(shp, 1.0, (0.5, 0.0, 0.0), 1.0)
    Wn = Wo.copy()
    c = tuple(4 if act[2 - a] else 0 for a in range(3))
    Wn[(var,) + c] *= 1.1
    tr, _ = detect(Wn, Wo, act)
    ix = interior(shp, act)
    expect = np.zeros(shp[1:])
    if var in (RHO, P):
        expect[c] = 1.0
    assert np.array_equal(tr[(var if var in (RHO, P) else RHO,) + ix], expect[ix])
    assert np.array_equal(tr[(SLOT,) + ix], expect[ix])


def test_nad_band_edges_relative_and_tolerance_zero():
    """Relative band [min - |min| tol, max + |max| tol] on a uniform field of 1:
    1 + tol/2 inside, 1 + 2 tol outside; with tol = 0 the band is exactly
    [min, max]: max itself inside, the next double above outside."""
    act = (1, 0, 0)
    shp = fv_shape(act, 6)
    Wo = prim_state(shp)
    for val, tol, flagged in [(1 + 0.5e-5, 1e-5, 0), (1 + 2e-5, 1e-5, 1), (1 - 2e-5, 1e-5, 1), (1 - 0.5e-5, 1e-5, 0),
                              (1.0, 0.0, 0), (np.nextafter(1.0, 2.0), 0.0, 1), (np.nextafter(1.0, 0.0), 0.0, 1)]:
        Wn = Wo.copy()
        Wn[RHO, 0, 0, 5] = val
        tr, _ = detect(Wn, Wo, act, nad_tol=tol)
        assert tr[RHO, 0, 0, 5] == flagged, (val, tol)


def test_nad_delta_band_differs_from_relative_band():
    """Old values 1 and 3 in the box: delta band (tol 0.5) is [0, 4],
    relative band [0.5, 4.5]; 4.1 is out of the first only, 0.2 out of both... 0.2
    is inside [0, 4] but outside [0.5, 4.5]."""
    act = (1, 0, 0)
    shp = fv_shape(act, 6)
    Wo = prim_state(shp)
    Wo[RHO, 0, 0, 6] = 3.0
    for val, delta_flag, rel_flag in [(4.1, 1, 0), (0.2, 0, 1), (3.9, 0, 0)]:
        Wn = Wo.copy()
        Wn[RHO, 0, 0, 5] = val
        td, _ = detect(Wn, Wo, act, nad_tol=0.5, nad_delta=1)
        trl, _ = detect(Wn, Wo, act, nad_tol=0.5, nad_delta=0)
        assert (td[RHO, 0, 0, 5], trl[RHO, 0, 0, 5]) == (delta_flag, rel_flag), val


@pytest.mark.parametrize("act", [(1, 1, 0), (1, 1, 1)])
def test_nad_diagonal_extremum_seen_by_moore_box_only(act):
    """The old maximum sits on a diagonal neighbour of the cell: the Moore box
    contains it (no flag), the face neighbourhood does not (flag)."""
    shp = fv_shape(act, 6)
    Wo = prim_state(shp)
    c = (4 if act[2] else 0, 4, 4)
    diag = (c[0] + (1 if act[2] else 0), c[1] + 1, c[2] + 1)
    Wo[(RHO,) + diag] = 2.0
    Wn = Wo.copy()
    Wn[(RHO,) + c] = 1.5
    tm, _ = detect(Wn, Wo, act, nad_moore=1)
    tf, _ = detect(Wn, Wo, act, nad_moore=0)
    assert tm[(RHO,) + c] == 0 and tf[(RHO,) + c] == 1
    ix = interior(shp, act)
    assert tf[(RHO,) + ix].sum() == 1 and tm[(RHO,) + ix].sum() == 0


@pytest.mark.parametrize("act", ACTS)
def test_nad_and_relax_write_cells_one_to_n_minus_two(act):
    """Every cell troubled (W_new = 2 W_old): flags on cells 1..N-2 of every
    active axis; cells 0 and N-1 keep what the array held (src/define.hpp,
    fv_for_cells_ngh). Unlimited variables are never written."""
    shp = fv_shape(act, 5)
    Wo = prim_state(shp)
    tr, _ = detect(2 * Wo, Wo, act, sentinel=7.0)
    ix = interior(shp, act)
    expect = np.full(shp[1:], 7.0)
    expect[ix] = 1.0
    for v in (RHO, P, SLOT):
        assert np.array_equal(tr[v], expect), v
    for v in (M1, M2, M3):
        assert np.all(tr[v] == 7.0)


# ---------------------------------------------------------------- SED and relax
def _axis_field(shp, act, f):
    """W[RHO] = f(x) + f(y) (+ f(z)) on unit-spaced cell centres."""
    W = prim_state(shp)
    idx = np.indices(shp[1:]).astype(float) + 0.5  # (z, y, x) centres
    W[RHO] = sum(f(idx[2 - d]) for d in range(3) if act[d]) + 5.0
    return W


@pytest.mark.parametrize("act", ACTS)
def test_sed_alpha_is_one_on_a_smooth_quadratic_and_unwritten_at_edges(act):
    """On a uniform grid a quadratic has d2u = 2, vL = vR = 2h, dv = h, so
    alpha = min(1, 2) = 1 -- also at its vertex (a smooth extremum).
    alpha is computed on cells 2..N-3 only; 0, 1, N-2, N-1 stay 0."""
    shp = fv_shape(act, 8)
    x0 = shp[3] / 2.0 + 0.3
    W = _axis_field(shp, act, lambda x: 0.02 * (x - x0) ** 2)
    _, al = detect(W, W, act, sed=1, p=3)
    ix2 = interior(shp, act, lo=2)
    for d in range(3):
        a = al[d][RHO]
        if not act[d]:
            assert np.all(a == 0)
            continue
        expect = np.zeros(shp[1:])
        expect[ix2] = 1.0
        assert np.array_equal(a, expect), d
        assert np.array_equal(al[d][P][ix2], np.ones(al[d][P][ix2].shape))  # constant pressure: dv = 0


def test_sed_step_keeps_alpha_below_one():
    act = (1, 0, 0)
    shp = fv_shape(act, 8)
    W = prim_state(shp)
    W[RHO, 0, 0, 6:] = 2.0  # step between cells 5 and 6
    _, al = detect(W, W, act, sed=1, p=3)
    a = al[0][RHO, 0, 0]
    assert a[5] < 1 and a[6] < 1
    assert a[2] == 1 and a[9] == 1  # far from the step: flat, alpha = 1


@pytest.mark.parametrize("act", ACTS)
def test_relax_clears_smooth_flags_except_next_to_unwritten_alpha(act):
    """W_old = W_new + 1 flags every cell by NAD; W_new is a smooth
    quadratic, so alpha = 1 wherever it is computed and relax clears a flag
    unless the +-1 neighbourhood along some axis reaches a never-written
    alpha (cells 0, 1, N-2, N-1): flags stay on cells 1, 2, N-3, N-2 of each
    active axis -- the first and last two active-frame rows."""
    shp = fv_shape(act, 8)
    Wn = _axis_field(shp, act, lambda x: 0.02 * (x - 3.7) ** 2)
    Wo = Wn.copy()
    Wo[RHO] += 1.0  # larger than any neighbour difference of W_new
    tr, _ = detect(Wn, Wo, act, sed=1, p=3, sentinel=0.0)
    keep = np.zeros(shp[1:], dtype=bool)
    for d in range(3):
        if not act[d]:
            continue
        n = shp[3 - d]
        idx = np.indices(shp[1:])[2 - d]
        keep |= (idx == 1) | (idx == 2) | (idx == n - 3) | (idx == n - 2)
    ix = interior(shp, act)
    expect = np.zeros(shp[1:])
    expect[ix] = keep[ix].astype(float)
    assert np.array_equal(tr[RHO], expect)
    assert np.array_equal(tr[SLOT], expect)
    # SED off (or p = 1): every NAD flag passes through
    t1, _ = detect(Wn, Wo, act, sed=1, p=1, sentinel=0.0)
    assert np.all(t1[(RHO,) + ix] == 1)


# ---------------------------------------------------------------- PAD
def test_pad_bounds():
    act = (1, 1, 0)
    shp = fv_shape(act, 6)
    W = prim_state(shp)
    W[RHO, 0, 3, 3] = 1e-11   # below rho_min
    W[P, 0, 4, 6] = 2e10      # above p_max
    W[RHO, 0, 5, 5] = 1e-10   # bounds are inclusive
    W[P, 0, 6, 4] = 1e10
    W[RHO, 0, 0, 0] = 1e-12   # outside the cells the check covers
    tr, _ = detect(W, W, act, pad=1, sentinel=0.0)
    expect = np.zeros(shp[1:])
    expect[0, 3, 3] = expect[0, 4, 6] = 1
    assert np.array_equal(tr[SLOT], expect)
    assert np.all(tr[RHO] == 0)  # PAD sets only the aggregate
    t0, _ = detect(W, W, act, pad=0, sentinel=0.0)
    assert np.all(t0[SLOT] == 0)


# ---------------------------------------------------------------- theta
W_BOX = {0: 1.0, 1: 0.75, 2: 0.5, 3: 0.375}


@pytest.mark.parametrize("act", ACTS)
def test_theta_box_weights_and_ring(act):
    """One troubled cell c: theta_tmp = weight by the number of nonzero
    offsets inside the 3^d box; theta = max(theta_tmp, 0.25 where any box
    neighbour has theta_tmp > 0), i.e. 0.25 on the cells at Chebyshev
    distance 2; both written on cells 1..N-2 only."""
    shp = fv_shape(act, 8)
    tr = np.zeros(shp)
    c = tuple(5 if act[2 - a] else 0 for a in range(3))
    tr[(SLOT,) + c] = 1.0
    tmp, th = blend(tr, act, sentinel=0.0)
    idx = np.indices(shp[1:])
    off = [idx[a] - c[a] for a in range(3)]
    cheb = np.max(np.abs(np.stack(off)), axis=0)
    nz = sum((o != 0).astype(int) for o in off)
    e_tmp = np.where(cheb <= 1, np.vectorize(W_BOX.get)(nz), 0.0)
    e_th = np.where(cheb == 2, 0.25, e_tmp)
    assert np.array_equal(tmp, e_tmp)
    assert np.array_equal(th, e_th)
    n_ring = {1: 2, 2: 16, 3: 98}[sum(act)]
    assert int((th == 0.25).sum()) == n_ring


def test_theta_max_of_overlapping_boxes_and_frame():
    """Two troubled cells two apart in x: the cell between them is a face
    neighbour of both (0.75); the trouble flag on a ghost cell (0) feeds the
    box of cell 1; theta_tmp is never written on cells 0 and N-1, so the ring
    there reads its zero initial value and theta stays what the array held."""
    act = (1, 0, 0)
    shp = fv_shape(act, 8)
    tr = np.zeros(shp)
    tr[SLOT, 0, 0, 4] = tr[SLOT, 0, 0, 6] = 1.0
    tr[SLOT, 0, 0, 0] = 1.0
    tmp, th = blend(tr, act, sentinel=9.0)
    assert list(tmp[0, 0]) == [0, 0.75, 0, 0.75, 1, 0.75, 1, 0.75, 0, 0, 0, 0]
    assert list(th[0, 0]) == [9.0, 0.75, 0.25, 0.75, 1, 0.75, 1, 0.75, 0.25, 0, 0, 9.0]


def test_theta_without_blending_is_the_aggregate_flag():
    act = (1, 1, 0)
    shp = fv_shape(act, 6)
    rng = np.random.default_rng(5)
    tr = np.zeros(shp)
    tr[SLOT] = (rng.random(shp[1:]) < 0.2).astype(float)
    _, th = blend(tr, act, blending=0, sentinel=9.0)
    assert np.array_equal(th, tr[SLOT])  # every cell, ghosts included, binary


# ---------------------------------------------------------------- MUSCL-Hancock + HLLC
def _written_faces(shp, act, d):
    """Faces the fallback writes along every axis: cells 2..N-2 of each active
    axis write their left face (fv_for_faces: the active cells 2..N-3 plus one)."""
    fshape = list(shp[1:])
    if act[d]:
        fshape[2 - d] += 1
    m = np.zeros(fshape, dtype=bool)
    sl = []
    for a in range(3):
        ax = 2 - a
        sl.append(slice(2, shp[1 + a] - 1) if act[ax] else slice(None))
    m[tuple(sl)] = True
    return m


@pytest.mark.parametrize("act", ACTS)
def test_fallback_theta_zero_keeps_sd_flux_and_face_range(act):
    shp = fv_shape(act, 6)
    rng = np.random.default_rng(11)
    W = prim_state(shp, 1.0, (0.2, -0.1, 0.3), 1.0)
    W[RHO] += 0.1 * rng.random(shp[1:])
    W[P] += 0.1 * rng.random(shp[1:])
    F0 = [rng.normal(size=(6,) + tuple(shp[1 + a] + (1 if (act[d] and 2 - a == d) else 0) for a in range(3)))
          for d in range(3)]
    F = fallback(W, np.zeros(shp[1:]), act, 0.01, F0)
    for d in range(3):
        assert np.array_equal(F[d], F0[d]), d  # F + 0 (F_HLLC - F) = F, bit for bit
    nanF = [np.full_like(f, np.nan) for f in F0]
    F = fallback(W, np.ones(shp[1:]), act, 0.01, nanF)
    for d in range(3):
        if not act[d]:
            assert np.all(np.isnan(F[d]))
            continue
        m = _written_faces(shp, act, d)
        # theta = 1: F + 1 (F_HLLC - F) is NaN where F was NaN; so count the
        # written faces on a zero-initialised array instead
        assert np.all(np.isnan(F[d]))
        Fz = fallback(W, np.ones(shp[1:]), act, 0.01, [np.zeros_like(f) for f in F0])[d]
        assert np.array_equal(np.any(Fz != 0, axis=0), m), d


@pytest.mark.parametrize("act", ACTS)
def test_fallback_uniform_state_gives_physical_flux(act):
    shp = fv_shape(act, 6)
    v = (0.4, -0.7, 0.25)
    W = prim_state(shp, 1.3, v, 2.1)
    F = fallback(W, np.ones(shp[1:]), act, 0.02)
    for d in range(3):
        if not act[d]:
            continue
        m = _written_faces(shp, act, d)
        phys = euler_flux(W[:, 0, 0, 0], d)
        got = F[d][:, m]
        assert np.allclose(got, phys[:, None], rtol=1e-14, atol=1e-15), d


def test_fallback_linear_profile_reconstructed_exactly_on_the_subgrid():
    """Linear primitives on the solver's non-uniform sub-grid (centres are
    face midpoints): both one-sided slopes equal the gradient, minmod keeps
    it, and with dt = 0 both face states equal the profile at the face, so
    the HLLC flux is the physical flux there."""
    act, N, p = (1, 0, 0), 4, 3
    shp = (6, 1, 1, N * (p + 1) + 4)
    geom = _geom(shp, act, real_p=p, N=(N, 1, 1))
    xc, xf = geom[0][0], geom[1][0]
    W = prim_state(shp, 0.0, (0, 0.3, 0), 0.0)
    W[RHO, 0, 0] = 1.0 + 0.5 * xc
    W[M1, 0, 0] = 0.2 - 0.3 * xc
    W[P, 0, 0] = 1.5 + 0.4 * xc
    F = fallback(W, np.ones(shp[1:]), act, 0.0, geom=geom)[0]
    for i in range(2, shp[3] - 1):
        x = xf[i]
        w = np.array([1.0 + 0.5 * x, 0.2 - 0.3 * x, 0.3, 0.0, 1.5 + 0.4 * x])
        assert np.allclose(F[:, 0, 0, i], euler_flux(w, 0), rtol=1e-13, atol=1e-14), i


@pytest.mark.parametrize("act,hs", [((1, 0, 0), (0.25, 1, 1)), ((1, 1, 0), (0.25, 0.25, 1)),
                                     ((1, 1, 1), (0.25, 0.25, 0.25)), ((1, 1, 0), (0.25, 0.125, 1)),
                                     ((1, 1, 1), (0.25, 0.125, 0.5))])
def test_hancock_step_advects_a_linear_density(act, hs):
    """rho linear along every active axis, uniform v and p: the primitive
    time derivative is drho/dt = -v . grad rho (velocity and pressure stay
    put), so with equal cell widths the Hancock-predicted face states are
    rho(x_face) - dt/2 (v . g) from both sides; a wrong sign, a wrong dt/h
    scale or a dropped transverse term moves them. The reference forms the
    time derivative from half-cell increments of every axis and divides by
    the width h_D of the face axis (src/hydro.cpp:674-679), so with unequal
    widths a transverse term carries h_t / h_D: that is what is expected."""
    shp = fv_shape(act, 6)
    hs = np.array(hs, dtype=float)
    C, Fc = _geom(shp, act)
    geom = ([c * hs[d] for d, c in enumerate(C)], [f * hs[d] for d, f in enumerate(Fc)])
    v, g, dt = np.array([0.6, -0.4, 0.3]), np.array([0.2, 0.64, -0.06]), 0.07
    W = prim_state(shp, 0.0, tuple(v), 1.2)
    idx = np.indices(shp[1:]).astype(float) + 0.5
    W[RHO] = 1.0 + sum(g[d] * hs[d] * idx[2 - d] for d in range(3) if act[d])
    F = fallback(W, np.ones(shp[1:]), act, dt, geom=geom)
    for d in range(3):
        if not act[d]:
            continue
        shift = -0.5 * dt * sum(v[a] * g[a] * hs[a] / hs[d] for a in range(3) if act[a])
        m = _written_faces(shp, act, d)
        fi = np.indices(m.shape).astype(float) + 0.5
        fi[2 - d] -= 0.5  # face coordinate along d
        rho_f = 1.0 + sum(g[a] * hs[a] * fi[2 - a] for a in range(3) if act[a]) + shift
        for pos in zip(*np.nonzero(m)):
            w = np.array([rho_f[pos], v[0], v[1], v[2], 1.2])
            assert np.allclose(F[d][(slice(None),) + pos], euler_flux(w, d), rtol=1e-12, atol=1e-14), (d, pos)


def test_fallback_face_weight_is_max_of_adjacent_theta():
    """theta = 0.5 on cell 6 only: faces 6 and 7 (its left and right) get
    F + 0.5 (F_HLLC - F), every other face keeps F."""
    act = (1, 0, 0)
    shp = fv_shape(act, 8)
    W = prim_state(shp, 1.0, (0.5, 0, 0), 1.0)
    th = np.zeros(shp[1:])
    th[0, 0, 6] = 0.5
    F0 = [np.full((6, 1, 1, shp[3] + 1), 3.0), np.zeros((6, 1, 1, shp[3])), np.zeros((6, 1, 1, shp[3]))]
    F = fallback(W, th, act, 0.01, F0)[0]
    phys = euler_flux(W[:, 0, 0, 0], 0)
    for i in range(shp[3] + 1):
        want = 3.0 + 0.5 * (phys - 3.0) if i in (6, 7) else np.full(6, 3.0)
        assert np.allclose(F[:, 0, 0, i], want, rtol=1e-15), i


# ---------------------------------------------------------------- boundaries
def _sd_trace_case(bc, dim, act):
    """A random physical state at the flux points of axis `dim` (one slice),
    its ghost traces filled, Euler fluxes and face LLF solves applied."""
    p, N = 2, 4
    n, m = p + 1, p + 2
    Ne = [N + 2 if act[d] else 1 for d in range(3)]
    npt = [(m if d == dim else n) if act[d] else 1 for d in range(3)]
    shape = (1, 6, Ne[2], Ne[1], Ne[0], npt[2], npt[1], npt[0])
    rng = np.random.default_rng(7 + bc + 3 * dim)
    U = np.zeros(shape)
    rho = rng.uniform(0.8, 1.2, shape[2:])
    vel = [rng.uniform(-0.5, 0.5, shape[2:]) for _ in range(3)]
    pr = rng.uniform(0.8, 1.2, shape[2:])
    U[0, 0] = rho
    for d in range(3):
        U[0, 1 + d] = rho * vel[d]
    U[0, 4] = pr / (GM - 1) + 0.5 * rho * sum(v * v for v in vel)
    cfg = op.OrcCfg.make(act=act, bc=tuple(bc if d == dim else 0 for d in range(3)), gamma=GM)
    s = op.ip(shape)
    op.orc().orc_sd_bc(ctypes.byref(cfg), op.dp(U), s, dim)
    F = np.zeros(shape)
    op.orc().orc_fluxes(ctypes.byref(cfg), op.dp(U), s, op.dp(F), dim)
    op.orc().orc_riemann(ctypes.byref(cfg), op.dp(U), op.dp(F), s, dim)
    return U, F, N, m


def _wall(arr, dim, elem, point):
    """Sub-array (t, var, rest...) at element `elem`, point `point` along `dim`."""
    idx = [slice(None)] * 8
    idx[4 - dim] = elem
    idx[7 - dim] = point
    return arr[tuple(idx)]


@pytest.mark.parametrize("act,dim", [((1, 0, 0), 0), ((1, 1, 0), 0), ((1, 1, 0), 1), ((1, 1, 1), 2)])
def test_reflective_wall_flux_is_a_wall(act, dim):
    """The wall Riemann problem sees the interior trace and its mirror: zero
    mass, transverse momentum and energy flux through both walls."""
    U, F, N, m = _sd_trace_case(2, dim, act)
    for elem, point in ((1, 0), (N, m - 1)):
        f = _wall(F, dim, elem, point)
        for v in (0, 4) + tuple(1 + d for d in range(3) if d != dim):
            assert np.all(f[:, v] == 0.0), (elem, v)
        assert np.all(f[:, 1 + dim] != 0.0)


@pytest.mark.parametrize("act,dim", [((1, 0, 0), 0), ((1, 1, 0), 1), ((1, 1, 1), 2)])
def test_gradfree_wall_flux_is_the_interior_physical_flux(act, dim):
    U, F, N, m = _sd_trace_case(1, dim, act)
    for elem, point in ((1, 0), (N, m - 1)):
        u = _wall(U, dim, elem, point)[0]
        f = _wall(F, dim, elem, point)[0]
        rho = u[0]
        vel = u[1:4] / rho
        pr = (u[4] - 0.5 * rho * (vel * vel).sum(axis=0)) * (GM - 1)
        want = np.zeros_like(u)
        want[0] = rho * vel[dim]
        for d in range(3):
            want[1 + d] = rho * vel[dim] * vel[d] + (pr if d == dim else 0)
        want[4] = (u[4] + pr) * vel[dim]
        assert np.allclose(f[:5], want[:5], rtol=1e-13, atol=1e-15), elem


@pytest.mark.parametrize("act,dim", [((1, 0, 0), 0), ((1, 1, 0), 1), ((1, 1, 1), 2)])
def test_periodic_walls_share_one_face_flux(act, dim):
    U, F, N, m = _sd_trace_case(0, dim, act)
    assert np.array_equal(_wall(F, dim, 1, 0), _wall(F, dim, N, m - 1))


@pytest.mark.parametrize("bc,expect", [
    # cells 2..5 active with values 10..13 (momentum x in variable 1)
    (0, [12, 13, 10, 11, 12, 13, 10, 11]),   # periodic: images one period away
    (1, [10, 11, 10, 11, 12, 13, 12, 13]),   # gradfree: the first / last two active cells, in order
    (2, [11, 10, 10, 11, 12, 13, 13, 12]),   # reflective: mirrored, normal momentum negated
])
def test_fv_ghost_layers_literal_1d(bc, expect):
    act = (1, 0, 0)
    U = np.full((6, 1, 1, 8), -99.0)
    for v in range(6):
        U[v, 0, 0, 2:6] = [10, 11, 12, 13]
    cfg = op.OrcCfg.make(act=act, bc=(bc, 0, 0))
    op.orc().orc_fv_bc(ctypes.byref(cfg), op.dp(U), op.ip(U.shape))
    for v in range(6):
        sign = -1 if (bc == 2 and v == 1) else 1
        want = np.array(expect, dtype=float)
        want[[0, 1, 6, 7]] *= sign
        assert np.array_equal(U[v, 0, 0], want), v


def test_fv_ghost_corners_2d_periodic_and_reflective():
    """x is filled before y, so a corner ghost is the image of an active cell
    through both walls (the Moore-box NAD and blending read corners)."""
    act = (1, 1, 0)
    base = np.full((6, 1, 8, 8), -99.0)
    vals = np.arange(16, dtype=float).reshape(4, 4) + 100
    base[:, 0, 2:6, 2:6] = vals
    for bc, corner00, corner77 in [(0, vals[2, 2], vals[1, 1]), (2, vals[1, 1], vals[2, 2])]:
        U = base.copy()
        cfg = op.OrcCfg.make(act=act, bc=(bc, bc, 0))
        op.orc().orc_fv_bc(ctypes.byref(cfg), op.dp(U), op.ip(U.shape))
        assert U[0, 0, 0, 0] == corner00 and U[0, 0, 7, 7] == corner77, bc
        if bc == 2:  # normal momenta negated once across their own wall, not across the other
            assert U[1, 0, 0, 0] == -corner00 and U[2, 0, 0, 0] == -corner00 and U[3, 0, 0, 0] == corner00


# ---------------------------------------------------------------- whole runs
def _sod_exact(x, t, x0=0.5, gm=GM, dL=1.0, pL=1.0, dR=0.125, pR=0.1):
    """Exact solution of the Sod Riemann problem (Newton iteration on the
    star pressure, Toro ch. 4), density at positions x."""
    cL, cR = math.sqrt(gm * pL / dL), math.sqrt(gm * pR / dR)

    def f(pk, dk, pk0, ck):
        if pk > pk0:  # shock
            A, B = 2 / ((gm + 1) * dk), (gm - 1) / (gm + 1) * pk0
            return (pk - pk0) * math.sqrt(A / (pk + B))
        return 2 * ck / (gm - 1) * ((pk / pk0) ** ((gm - 1) / (2 * gm)) - 1)

    ps = 0.5 * (pL + pR)
    for _ in range(100):
        g = f(ps, dL, pL, cL) + f(ps, dR, pR, cR)
        h = 1e-8 * ps
        dg = (f(ps + h, dL, pL, cL) + f(ps + h, dR, pR, cR) - g) / h
        ps -= g / dg
    us = 0.5 * (f(ps, dR, pR, cR) - f(ps, dL, pL, cL))
    dsL = dL * (ps / pL) ** (1 / gm)  # left rarefaction
    dsR = dR * ((ps / pR + (gm - 1) / (gm + 1)) / ((gm - 1) / (gm + 1) * ps / pR + 1))  # right shock
    sR = us * dsR / (dsR - dR)  # shock speed from mass conservation (uR = 0)
    csL = cL * (ps / pL) ** ((gm - 1) / (2 * gm))
    xi = (x - x0) / t
    rho = np.empty_like(x)
    head, tail = -cL, us - csL
    for k, s in enumerate(xi):
        if s < head:
            rho[k] = dL
        elif s < tail:
            c = 2 / (gm + 1) * (cL + (gm - 1) / 2 * (0.0 - s))
            rho[k] = dL * (c / cL) ** (2 / (gm - 1))
        elif s < us:
            rho[k] = dsL
        elif s < sR:
            rho[k] = dsR
        else:
            rho[k] = dR
    return rho


def test_muscl_hancock_hllc_alone_converges_to_exact_sod(tmp_path):
    """theta = 1 on every cell (a negative NAD tolerance flags every cell,
    SED off): the update is the MUSCL-Hancock + HLLC scheme on the sub-grid.
    L1(rho) against the exact solution at t = 0.2 falls by more than 0.7 per
    doubling of the grid (first order at the discontinuities)."""
    errs = []
    for N in (16, 32, 64):
        d = tmp_path / f"n{N}"
        op.run_oracle(["-i", os.path.join(op.INPUTS, "sod.athinput"), f"mesh/nx1={N}", "mesh/nx2=1", "mesh/nx3=1",
                       "mesh/p=1", "fallback/tolerance=-1", "fallback/SED=false", "time/tlim=0.2",
                       "output/dt=0.2"], str(d))
        g = op.grid(str(d))
        i = op.output_indices(str(d))[-1]
        rho = op.load_sd(str(d), g, i, 0).ravel()
        faces = g["faces"][0]
        xc = 0.5 * (faces[1:] + faces[:-1])
        # exact cell average by fine quadrature
        sub = np.linspace(0, 1, 41)[:-1] + 1 / 80
        ex = np.array([_sod_exact(a + (b - a) * sub, 0.2).mean() for a, b in zip(faces[:-1], faces[1:])])
        assert len(ex) == len(rho)
        errs.append(float(np.sum(np.abs(rho - ex) * np.diff(faces))))
        assert np.all(np.isfinite(rho)) and xc[0] > 0
    assert errs[1] < 0.7 * errs[0] and errs[2] < 0.7 * errs[1], errs
    assert errs[2] < 0.01, errs


@pytest.mark.parametrize("force", [False, True])
def test_reflective_walls_equal_periodic_for_a_centred_blast(tmp_path, force):
    """A blast centred in the box is symmetric about x = 0.5 and, being
    periodic, about the walls: the periodic run's walls are symmetry planes,
    so the reflective run (SD traces and FV ghosts mirrored) must give the
    same fields and the same troubled cells (within round-off)."""
    base = ["-i", os.path.join(op.INPUTS, "sine_wave.athinput"), "problem/problem=spherical_blast",
            "mesh/nx1=6", "mesh/nx2=6", "mesh/nx3=1", "job/fallback=true", "hydro/gamma=1.6666666666667",
            "time/tlim=0.06", "output/dt=0.06"]
    if force:
        base += ["fallback/tolerance=-1", "fallback/SED=false"]
    runs = {}
    for name, bc in (("per", "periodic"), ("ref", "reflective")):
        d = tmp_path / name
        op.run_oracle(base + [f"mesh/x1_bc={bc}", f"mesh/x2_bc={bc}"], str(d))
        runs[name] = str(d)
    g = op.grid(runs["per"])
    i = op.output_indices(runs["per"])[-1]
    for v in range(5):
        a, b = op.load_sd(runs["per"], g, i, v), op.load_sd(runs["ref"], g, i, v)
        scale = max(np.abs(a).max(), 1e-300)
        assert np.abs(a - b).max() / scale < 1e-10, v
    ta = np.fromfile(os.path.join(runs["per"], f"troubles_N28_{i}_0.dat"))  # FV cells incl. ghosts
    tb = np.fromfile(os.path.join(runs["ref"], f"troubles_N28_{i}_0.dat"))
    assert np.array_equal(ta, tb)
    assert ta.sum() > 0  # the detector fired
