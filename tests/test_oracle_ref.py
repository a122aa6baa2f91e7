"""Pins the oracle's ModuleFunctions restatement to the reference itself.

The reference's ModuleFunctions.F90 is compiled unmodified from source into
oracle/_ref (make -C oracle ref); these tests call both on identical seeded
inputs and require bit-identical results:

* THOMASZ_NewType2 (ModuleFunctions.F90:4026-4123), including the silent
  zero-pivot skip (MF:4092-4098) with identical VEC scratch contents;
* THOMAS_3D (its T_THOMAS form, MF:3667-3870) along j and along i, with
  identical VEC scratch contents afterwards;
* ComputeAdvectionFace (MF:10702-10894) for every method, limiter, flow sign
  and NearBoundary/Upwind2 combination the reference accepts;
* OrlanskiCelerity2D (MF:4129-4490) as AdvectionDiffusion calls it
  (AD:5506-5568) at every boundary cell of a grid, corners included, both
  flow signs, with and without celerity clipping: identical new values and
  identical updates of the new and old fields;
* ComputeAdvection1D_V2 (MF:10534-10591) and
  ComputeAdvection1D_TVD_SuperBee_2 (MF:10642-10699) on random lines with
  random compute masks (which exercises every NearBoundary branch and the
  partial writes of the _2 builder).
"""
import ctypes as ct

import numpy as np
import pytest

from mohid_amd import constants as C
from oracle import oracle as ora

pytestmark = pytest.mark.ref

dp = ct.POINTER(ct.c_double)
ip = ct.POINTER(ct.c_int32)


def P(a, kind=dp):
    return a.ctypes.data_as(kind)


def _thomas_inputs(rng, ni, nj, nk, zero_pivot=False):
    size = (0, ni + 1, 0, nj + 1, 0, nk + 1)
    shape = (ni + 2, nj + 2, nk + 2)
    D = np.asfortranarray(-rng.uniform(0, 1, shape))
    F = np.asfortranarray(-rng.uniform(0, 1, shape))
    E = np.asfortranarray(1.0 + rng.uniform(0, 1, shape) - D - F)
    TI = np.asfortranarray(rng.normal(0, 3, shape))
    WP = np.asfortranarray((rng.uniform(size=shape) < 0.8).astype(np.int32))
    if zero_pivot:
        # make row 3 of column (1,1) singular after elimination: E = -D*W(2)
        # is hard to hit exactly, so use W(2)=0 (F(2)=0) and E(3)=0
        F[1, 1, 2] = 0.0
        E[1, 1, 3] = 0.0
        WP[1, 1, nk] = 1
    return size, (1, ni, 1, nj, 1, nk), D, E, F, TI, WP


@pytest.mark.parametrize("seed,zero_pivot", [(1, False), (2, False), (3, True)])
def test_thomasz_bitwise(seed, zero_pivot):
    rng = np.random.default_rng(seed)
    ni, nj, nk = 7, 5, 9
    size, ws, D, E, F, TI, WP = _thomas_inputs(rng, ni, nj, nk, zero_pivot)
    lo, hi = 0, max(ni, nj, nk) + 1
    r_ora = np.asfortranarray(rng.normal(size=D.shape))
    r_ref = r_ora.copy(order="F")
    w0 = rng.normal(size=hi - lo + 1)
    g0 = rng.normal(size=hi - lo + 1)
    w_ora, g_ora = w0.copy(), g0.copy()
    w_ref, g_ref = w0.copy(), g0.copy()
    S = ora.Size3D(*size)
    W = ora.Size3D(*ws)
    ora.lib().ora_thomasz(ct.byref(S), ct.byref(W), P(D), P(E), P(F), P(TI), P(WP, ip),
                          P(r_ora), P(w_ora), P(g_ora))
    sz = np.array(size, dtype=np.int32)
    wz = np.array(ws, dtype=np.int32)
    ora.ref().ref_thomasz(P(sz, ip), P(wz, ip), lo, hi, P(D), P(E), P(F), P(TI), P(WP, ip),
                          P(r_ref), P(w_ref), P(g_ref))
    assert np.array_equal(r_ora, r_ref, equal_nan=True)
    assert np.array_equal(w_ora, w_ref) and np.array_equal(g_ora, g_ref)


@pytest.mark.parametrize("seed,di,dj", [(11, 0, 1), (12, 1, 0), (13, 0, 1), (14, 1, 0)])
def test_thomas3d_bitwise(seed, di, dj):
    rng = np.random.default_rng(seed)
    ni, nj, nk = 9, 6, 4
    size, ws, D, E, F, TI, _ = _thomas_inputs(rng, ni, nj, nk)
    # the reference's calls: lines over the WorkSize, rows JImin..JImax+1
    # (AD:4194, 4233), and a sub-range to show the bounds are honoured
    if di == 0:
        lim = (1, ni, 1, nj, 1, nk) if seed < 13 else (2, ni - 1, 2, nj - 2, 2, nk)
    else:
        lim = (1, nj, 1, ni, 1, nk) if seed < 13 else (2, nj - 1, 1, ni - 3, 1, nk - 1)
    lo, hi = 0, max(ni, nj, nk) + 1
    r_ora = np.asfortranarray(rng.normal(size=D.shape))
    r_ref = r_ora.copy(order="F")
    w0, g0 = rng.normal(size=hi - lo + 1), rng.normal(size=hi - lo + 1)
    w_ora, g_ora, w_ref, g_ref = w0.copy(), g0.copy(), w0.copy(), g0.copy()
    S = ora.Size3D(*size)
    rc = ora.lib().ora_thomas3d(ct.byref(S), *lim, di, dj, P(D), P(E), P(F), P(TI), P(r_ora),
                                P(w_ora), P(g_ora), None)
    assert rc == 0
    sz = np.array(size, dtype=np.int32)
    lz = np.array(lim, dtype=np.int32)
    ora.ref().ref_thomas3d(P(sz, ip), P(lz, ip), di, dj, lo, hi, P(D), P(E), P(F), P(TI),
                           P(r_ref), P(w_ref), P(g_ref))
    assert np.array_equal(r_ora, r_ref)
    assert np.array_equal(w_ora, w_ref) and np.array_equal(g_ora, g_ref)
    # rows outside the lines are left alone
    touched = np.zeros(D.shape, dtype=bool)
    a0, a1, b0, b1, k0, k1 = lim
    if di == 0:
        touched[a0:a1 + 1, b0:b1 + 2, k0:k1 + 1] = True
    else:
        touched[b0:b1 + 2, a0:a1 + 1, k0:k1 + 1] = True
    r0 = np.asfortranarray(np.random.default_rng(seed).normal(size=D.shape))
    assert not np.array_equal(r_ora[touched], r0[touched])


@pytest.mark.parametrize("di,dj", [(0, 1), (1, 0)])
def test_thomas3d_matches_dense_solve(di, dj):
    """Each line of THOMAS_3D equals a dense solve (numpy) of its tridiagonal
    system, rows JImin..JImax+1 with D of the first and F of the last row
    dropped."""
    rng = np.random.default_rng(21 + di)
    ni, nj, nk = 8, 7, 3
    size, ws, D, E, F, TI, _ = _thomas_inputs(rng, ni, nj, nk)
    lim = (1, ni, 1, nj, 1, nk) if di == 0 else (1, nj, 1, ni, 1, nk)
    res = np.asfortranarray(np.zeros(D.shape))
    hi = max(ni, nj, nk) + 1
    w, g = np.zeros(hi + 1), np.zeros(hi + 1)
    S = ora.Size3D(*size)
    assert ora.lib().ora_thomas3d(ct.byref(S), *lim, di, dj, P(D), P(E), P(F), P(TI), P(res),
                                  P(w), P(g), None) == 0
    for k in range(1, nk + 1):
        for ij in range(lim[0], lim[1] + 1):
            rows = range(lim[2], lim[3] + 2)
            sel = (lambda r: (ij, r, k)) if di == 0 else (lambda r: (r, ij, k))
            n = len(rows)
            A = np.zeros((n, n))
            b = np.zeros(n)
            for m, r in enumerate(rows):
                A[m, m] = E[sel(r)]
                if m > 0:
                    A[m, m - 1] = D[sel(r)]
                if m < n - 1:
                    A[m, m + 1] = F[sel(r)]
                b[m] = TI[sel(r)]
            x = np.linalg.solve(A, b)
            got = np.array([res[sel(r)] for r in rows])
            np.testing.assert_allclose(got, x, rtol=1e-11, atol=1e-12)


def test_thomas3d_zero_pivot_reported():
    """A zero pivot is reported with its cell (the reference stops, MF:3788)."""
    rng = np.random.default_rng(5)
    ni, nj, nk = 4, 5, 2
    size, ws, D, E, F, TI, _ = _thomas_inputs(rng, ni, nj, nk)
    F[2, 2, 1] = 0.0        # W(2) = 0 on line (i=2, k=1), so AUX(3) = E(3)
    E[2, 3, 1] = 0.0
    res = np.asfortranarray(np.zeros(D.shape))
    w, g = np.zeros(8), np.zeros(8)
    where = np.zeros(3, dtype=np.int32)
    S = ora.Size3D(*size)
    rc = ora.lib().ora_thomas3d(ct.byref(S), 1, ni, 1, nj, 1, nk, 0, 1, P(D), P(E), P(F),
                                P(TI), P(res), P(w), P(g), P(where, ip))
    assert rc == ora.ERR_PIVOT
    assert tuple(where) == (2, 3, 1)


def test_thomasz_matches_dense_solve():
    """THOMASZ of a water column equals a dense LU solve (numpy) of the same
    tridiagonal system, rows 1..KUB+1 (the ghost top row included)."""
    rng = np.random.default_rng(7)
    ni, nj, nk = 3, 2, 12
    size, ws, D, E, F, TI, WP = _thomas_inputs(rng, ni, nj, nk)
    WP[...] = 1
    res = np.zeros_like(D, order="F")
    w = np.zeros(nk + 3); g = np.zeros(nk + 3)
    ora.lib().ora_thomasz(ct.byref(ora.Size3D(*size)), ct.byref(ora.Size3D(*ws)), P(D), P(E),
                          P(F), P(TI), P(WP, ip), P(res), P(w), P(g))
    for i in range(1, ni + 1):
        for j in range(1, nj + 1):
            n = nk + 1
            A = np.zeros((n, n))
            for r in range(n):
                k = r + 1
                A[r, r] = E[i, j, k]
                if r > 0:
                    A[r, r - 1] = D[i, j, k]
                if r < n - 1:
                    A[r, r + 1] = F[i, j, k]
            x = np.linalg.solve(A, TI[i, j, 1:n + 1])
            np.testing.assert_allclose(res[i, j, 1:n + 1], x, rtol=1e-13, atol=1e-13)


METHODS = [C.UPWIND_ORDER1, C.UPWIND_ORDER2, C.UPWIND_ORDER3, C.P2_TVD, C.CENTRAL_DIF,
           C.LEAP_FROG]
LIMITERS = [C.MINMOD, C.VANLEER, C.MUSCL, C.SUPERBEE, C.PDM]


def _reference_accepts(method, nearb, upwind2):
    if method in (C.UPWIND_ORDER1, C.CENTRAL_DIF, C.LEAP_FROG):
        return True
    return (not nearb) or upwind2


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("tvd", LIMITERS)
def test_advection_face_bitwise(method, tvd):
    rng = np.random.default_rng(100 * method + tvd)
    n_checked = 0
    for trial in range(300):
        prop = rng.normal(5.0, 2.0, 4)
        if trial % 10 == 0:
            prop[:] = 3.25                      # uniform: dC clamp branch
        if trial % 10 == 1:
            prop = 1.0 + 0.5 * np.arange(4)     # linear: r = 1 on a uniform grid
        vol = rng.uniform(1e3, 5e3, 4)
        du = rng.uniform(100.0, 600.0, 4)
        if trial % 10 == 1:
            du[:] = 250.0
        q = rng.normal(0.0, 50.0)
        if trial % 17 == 0:
            q = 0.0
        dt = float(rng.uniform(10, 120))
        nearb = int(rng.uniform() < 0.3)
        up2 = int(rng.uniform() < 0.6)
        if not _reference_accepts(method, nearb, up2):
            c = np.zeros(4)
            assert ora.lib().ora_advection_face(P(prop), P(vol), P(du), dt, q, 1.5, method, tvd,
                                                nearb, up2, P(c)) == ora_err_method()
            continue
        c_ora = np.full(4, np.nan)
        c_ref = np.full(4, np.nan)
        rc = ora.lib().ora_advection_face(P(prop), P(vol), P(du), dt, q, 1.5, method, tvd,
                                          nearb, up2, P(c_ora))
        assert rc == 0
        ora.ref().ref_advection_face(P(prop), P(vol), P(du), dt, q, 1.5, method, tvd, nearb,
                                     up2, P(c_ref))
        assert np.array_equal(c_ora, c_ref), (trial, prop, q, nearb, up2, c_ora, c_ref)
        n_checked += 1
    assert n_checked > 100


def ora_err_method():
    return 201


def _random_line(rng, n, p_open=0.8):
    du = rng.uniform(100.0, 600.0, n)
    prop = np.exp(rng.normal(0.0, 1.0, n))
    q = rng.normal(0.0, 40.0, n)
    q[rng.uniform(size=n) < 0.05] = 0.0
    v = rng.uniform(1e4, 1e5, n)
    cp = (rng.uniform(size=n) < p_open).astype(np.int32)
    cp[0] = 0  # halo: the reference would read before the line otherwise
    cp[-1] = 0
    return du, prop, q, v, cp


@pytest.mark.parametrize("method,tvd", [(C.UPWIND_ORDER1, C.SUPERBEE), (C.P2_TVD, C.SUPERBEE),
                                        (C.P2_TVD, C.MINMOD), (C.P2_TVD, C.VANLEER),
                                        (C.P2_TVD, C.MUSCL), (C.P2_TVD, C.PDM),
                                        (C.UPWIND_ORDER2, C.SUPERBEE),
                                        (C.UPWIND_ORDER3, C.SUPERBEE),
                                        (C.CENTRAL_DIF, C.SUPERBEE)])
def test_advection1d_v2_bitwise(method, tvd):
    rng = np.random.default_rng(method * 10 + tvd)
    for trial in range(20):
        n = int(rng.integers(6, 40))
        du, prop, q, v, cp = _random_line(rng, n)
        outs_ora = [np.full(n, -7.5) for _ in range(4)]
        outs_ref = [np.full(n, -7.5) for _ in range(4)]
        rc = ora.lib().ora_advection1d_v2(n, 2, n - 1, 30.0, P(du), P(prop), P(q), P(v),
                                          P(cp, ip), *[P(o) for o in outs_ora], method, tvd,
                                          1.5, 1)
        assert rc == 0
        ora.ref().ref_advection1d_v2(n, 2, n - 1, 30.0, P(du), P(prop), P(q), P(v), P(cp, ip),
                                     *[P(o) for o in outs_ref], method, tvd, 1.5, 1)
        for a, b in zip(outs_ora, outs_ref):
            assert np.array_equal(a, b)


def test_advection1d_tvd_superbee_2_bitwise():
    rng = np.random.default_rng(42)
    for trial in range(40):
        n = int(rng.integers(6, 40))
        du, prop, q, v, cp = _random_line(rng, n)
        d_ora, e_ora = np.full(n, 11.0), np.full(n, -3.0)
        d_ref, e_ref = d_ora.copy(), e_ora.copy()
        ora.lib().ora_advection1d_tvd_superbee_2(n, 2, n - 1, 60.0, P(du), P(prop), P(q), P(v),
                                                 P(cp, ip), P(d_ora), P(e_ora))
        ora.ref().ref_advection1d_tvd_superbee_2(n, 2, n - 1, 60.0, P(du), P(prop), P(q), P(v),
                                                 P(cp, ip), P(d_ref), P(e_ref))
        assert np.array_equal(d_ora, d_ref) and np.array_equal(e_ora, e_ref)


def orlanski2d_ref_diff(seed, limit):
    """Largest |oracle - reference| over the calls and the final fields, as
    test_orlanski2d_bitwise makes them."""
    try:
        test_orlanski2d_bitwise(seed, limit)
    except AssertionError:
        return float("inf")
    return 0.0


@pytest.mark.parametrize("seed,limit", [(1, 1e9), (2, 1e-3), (3, 0.05)])
def test_orlanski2d_bitwise(seed, limit):
    """Every edge cell of a 9 x 8 x 4 grid, in the reference's order, both
    codes on their own copies of the new/old fields; the mask has closed
    cells so that both one-sided cases occur."""
    rng = np.random.default_rng(seed)
    ni, nj, nk = 9, 8, 4
    sz = np.array([0, ni + 1, 0, nj + 1, 0, nk + 1], dtype=np.int32)
    shape = (ni + 2, nj + 2, nk + 2)
    newf = np.asfortranarray(rng.uniform(1.0, 3.0, shape))
    oldf = np.asfortranarray(newf + rng.normal(0, 0.05, shape))
    ref = np.asfortranarray(rng.uniform(1.0, 3.0, shape))
    cp = np.asfortranarray((rng.uniform(size=shape) < 0.85).astype(np.int32))
    a_new, a_old = newf.copy(order="F"), oldf.copy(order="F")
    b_new, b_old = newf.copy(order="F"), oldf.copy(order="F")
    size = ora.Size3D(*[int(v) for v in sz])
    n = 0
    for j in range(1, nj + 1):
        for i in range(1, ni + 1):
            if not (i in (1, ni) or j in (1, nj)):
                continue
            di, dj = (1, 0) if i in (1, ni) else (0, 1)
            en = 0 if (i == 1 or j == 1) else 1
            ie, je = (i + di, j + dj) if en else (i - di, j - dj)
            for k in range(1, nk + 1):
                flow = float(rng.normal(0, 0.3))
                got = ora.lib().ora_orlanski2d(ct.byref(size), P(a_new), P(a_old), P(ref),
                                               P(cp, ip), 1, ni, 1, nj, di, dj, ie, je, k, en,
                                               limit, 60.0, flow)
                want = ct.c_double()
                ora.ref().ref_orlanski2d(P(sz, ip), P(b_new), P(b_old), P(ref), P(cp, ip), 1,
                                         ni, 1, nj, di, dj, ie, je, k, en, limit, 60.0, flow,
                                         ct.byref(want))
                assert got == want.value or (np.isnan(got) and np.isnan(want.value)), \
                    (i, j, k, got, want.value)
                n += 1
    assert np.array_equal(a_new, b_new) and np.array_equal(a_old, b_old)
    assert n > 100
