"""Pins the oracle's initial conditions (row A2) function by function.

Each reference case is evaluated here in numpy from its closed form, as a
whole array over a non-cubic grid, and compared with the oracle's array:

* vlp4d_mpi (VMO/Init.cpp:152-277): CYL02 f = 4 h(x) h(vx), CYL05
  f = 4 h(y) h(vy) written only for ix < N0-1, iy < N1-1, ivx < N2-1 (its
  loops use '<' there, VMO/Init.cpp:196-199, so the last index of each of
  those dimensions stays zero), SLD10, TSI20 (two drifting Maxwellians at
  +-2.4 with vthy = 0.5);
* the Lagrange miniapp (miniapps/vlp4d/openmp/init.cpp:144-285): the same
  CYL02/SLD10, CYL05 over the full range, and its own TSI20 (the
  xi = 0.90 bump, unrelated to the vlp4d_mpi one);
* h(s) is the bump cos(pi/2 (s - cc)/rc) on [cc - rc, cc + rc],
  -cos(pi/2 (s + cc)/rc) on [-cc - rc, -cc + rc], 0 elsewhere, with
  cc = 3/16, rc = 1/8 (VMO/Init.cpp:156-157).

Grid point i of dimension d is minPhy_d + i dx_d with dx_d = (max-min)/N_d
(VMO/Init.cpp:80-82).  The bar is 4e-16 of the array's max-norm: the only
difference allowed is the last bit of exp/cos/sin between numpy and libm.
The halo cells of the oracle's (N+6)^4 array must stay exactly zero.  The
density integral of each case is checked against its own closed form where
one exists (SLD10, TSI20), which pins the velocity normalisation
independently of the pointwise formula.
"""
import math

import numpy as np
import pytest

import oracle

CC, RC = 0.50 * (6.0 / 16.0), 0.50 * (4.0 / 16.0)


def bump(s):
    s = np.asarray(s, dtype=np.float64)
    out = np.zeros_like(s)
    a = (s <= CC + RC) & (s >= CC - RC)
    b = ~a & (s <= -CC + RC) & (s >= -CC - RC)
    out[a] = np.cos(0.5 * math.pi * ((s[a] - CC) / RC))
    out[b] = -np.cos(0.5 * math.pi * ((s[b] + CC) / RC))
    return out


def axes(c):
    """Grid coordinates (x, y, vx, vy) as the reference forms them."""
    return [c.minPhy[d] + np.arange(c.nxmax[d], dtype=np.float64) * c.dx[d] for d in range(4)]


def bcast(v, d):
    """Put a 1-D array along dimension d of a [vy][vx][y][x] array."""
    shape = [1, 1, 1, 1]
    shape[3 - d] = v.size
    return v.reshape(shape)


def closed_form(c, idcase, lagrange=False):
    x, y, vx, vy = axes(c)
    X, Y, VX, VY = bcast(x, 0), bcast(y, 1), bcast(vx, 2), bcast(vy, 3)
    shape = tuple(c.nxmax[d] for d in (3, 2, 1, 0))
    if idcase == 2:
        f = 4.0 * bump(X) * bump(VX) * np.ones(shape)
    elif idcase == 5:
        f = 4.0 * bump(Y) * bump(VY) * np.ones(shape)
        if not lagrange:
            f[:, -1, :, :] = 0.0   # ivx = N2-1
            f[:, :, -1, :] = 0.0   # iy = N1-1
            f[:, :, :, -1] = 0.0   # ix = N0-1
    elif idcase == 10:
        f = (1.0 / (2 * math.pi)) * np.exp(-0.5 * (VX * VX + VY * VY)) * (1 + 0.05 * (np.cos(0.5 * X) * np.cos(0.5 * Y)))
    elif idcase == 20 and not lagrange:
        vd, vthx, vthy, ax, ay, k = 2.4, 1.0, 0.5, 0.05, 0.25, 0.2
        f = (1.0 / (4 * math.pi * vthx * vthy) * (1 - ax * np.sin(k * X) - ay * np.sin(k * Y))
             * (np.exp(-0.5 * (((VX - vd) / vthx) ** 2 + (VY / vthy) ** 2))
                + np.exp(-0.5 * (((VX + vd) / vthx) ** 2 + (VY / vthy) ** 2))))
    elif idcase == 20 and lagrange:
        xi, alpha = 0.90, 0.05
        f = ((1 + alpha * np.cos(0.5 * X)) / (2 * math.pi) * ((2 - 2 * xi) / (3 - 2 * xi))
             * (1 + 0.5 * VX * VX / (1 - xi)) * np.exp(-0.5 * (VX * VX + VY * VY)) * np.ones((1, 1, Y.size, 1)))
    else:
        raise ValueError(idcase)
    return np.broadcast_to(f, shape)


def cfg_for(idcase, n=(14, 11, 18, 13)):
    if idcase in (2, 5):
        # bumps live in [-0.3125, 0.3125]; boxes straddle them unevenly
        mn, mx = [-0.47, -0.41, -0.52, -0.38], [0.53, 0.44, 0.39, 0.33]
    elif idcase == 10:
        mn, mx = [0.0, 0.0, -6.0, -7.0], [4 * math.pi, 12.56637112, 6.0, 7.0]
    else:
        mn, mx = [0.0, 0.0, -8.0, -5.0], [10 * math.pi, 10 * math.pi, 8.0, 5.0]
    return oracle.make_config(n, mn, mx, idcase, 0.01, 1)


def assert_pinned(got, ref):
    scale = np.max(np.abs(ref))
    assert scale > 0
    err = np.max(np.abs(got - ref))
    assert err <= 4e-16 * scale, err / scale


@pytest.mark.parametrize("idcase", [2, 5, 10, 20])
def test_spline_init_closed_form(orc, idcase):
    c = cfg_for(idcase)
    h = orc.init_case(c)
    got = orc.interior_of(h)
    ref = closed_form(c, idcase)
    assert_pinned(got, ref)
    if idcase in (2, 5):
        assert np.count_nonzero(ref) > 20   # the grid really crosses the bumps
    # everything outside the interior is untouched (zero on entry, VMO/Init.cpp:370-371)
    rest = h.copy()
    orc.interior_of(rest)[...] = 0.0
    assert not np.any(rest)


def test_cyl05_leaves_last_indices_zero(orc):
    c = cfg_for(5)
    f = orc.interior_of(orc.init_case(c))
    assert not np.any(f[:, -1, :, :]) and not np.any(f[:, :, -1, :]) and not np.any(f[:, :, :, -1])
    # the rest, including the last vy index (its loop uses '<='), is the full bump
    full = closed_form(c, 5, lagrange=True)
    assert np.any(full[-1, :-1, :-1, :-1])
    assert_pinned(f[:, :-1, :-1, :-1], full[:, :-1, :-1, :-1])


@pytest.mark.parametrize("idcase", [2, 5, 10, 20])
def test_lagrange_init_closed_form(orc, idcase):
    c = cfg_for(idcase)
    got = orc.lag_init_case(c)
    assert_pinned(got, closed_form(c, idcase, lagrange=True))


def test_sld10_density_closed_form(orc):
    """rho(x, y) = dvx dvy sum_v f = (1 + 0.05 cos(x/2) cos(y/2)) S_x S_y with
    S_d = sum_i exp(-v_i^2/2) dv_d / sqrt(2 pi), which for a box of +-6 or
    more is 1 to 1e-8 (trapezoid rule on a Gaussian)."""
    c = cfg_for(10, n=(12, 10, 32, 30))
    rho = orc.field_rho(c, orc.init_case(c))
    x, y, vx, vy = axes(c)
    sx = np.sum(np.exp(-0.5 * vx ** 2)) * c.dx[2] / math.sqrt(2 * math.pi)
    sy = np.sum(np.exp(-0.5 * vy ** 2)) * c.dx[3] / math.sqrt(2 * math.pi)
    assert abs(sx - 1) < 1e-8 and abs(sy - 1) < 1e-8
    ref = (1 + 0.05 * np.cos(0.5 * x)[None, :] * np.cos(0.5 * y)[:, None]) * sx * sy
    np.testing.assert_allclose(rho, ref, rtol=1e-14, atol=0)


def test_tsi20_density_closed_form(orc):
    """Each beam integrates to 2 pi vthx vthy, so rho = (1 - 0.05 sin(0.2x) -
    0.25 sin(0.2y)) to trapezoid accuracy."""
    c = cfg_for(20, n=(12, 10, 48, 40))
    rho = orc.field_rho(c, orc.init_case(c))
    x, y, _, _ = axes(c)
    ref = 1 - 0.05 * np.sin(0.2 * x)[None, :] - 0.25 * np.sin(0.2 * y)[:, None]
    np.testing.assert_allclose(rho, ref, rtol=1e-7, atol=0)


def test_lagrange_tsi20_density_closed_form(orc):
    """The xi bump's velocity integral: int (1 + vx^2 / (2 (1 - xi))) e^{-v^2/2} dv
    = 2 pi (1 + 1 / (2 (1 - xi))) = 2 pi (3 - 2 xi) / (2 - 2 xi), so
    rho = 1 + 0.05 cos(x/2) in the continuum; the box's vy = +-5 cuts
    erfc(5/sqrt2) = 5.7e-7 of the vy Gaussian, hence the 1e-6 bar."""
    c = cfg_for(20, n=(12, 10, 64, 48))
    f = orc.lag_init_case(c)
    rho = orc.lag_field_rho(c, f)
    x = axes(c)[0]
    np.testing.assert_allclose(rho, np.broadcast_to(1 + 0.05 * np.cos(0.5 * x)[None, :], rho.shape), rtol=1e-6)


# ---- extension cases (no reference counterpart): the workload configurations'
# strong-Landau (30) and two-stream (31) initial conditions ------------------

def ext_closed_form(c, idcase):
    x, y, vx, vy = axes(c)
    X, Y, VX, VY = bcast(x, 0), bcast(y, 1), bcast(vx, 2), bcast(vy, 3)
    shape = tuple(c.nxmax[d] for d in (3, 2, 1, 0))
    if idcase == 30:
        f = np.exp(-0.5 * (VX ** 2 + VY ** 2)) / (2 * math.pi) * (1 + 0.5 * np.cos(0.5 * X) + 0.5 * np.cos(0.5 * Y))
    else:
        f = ((1 + 1e-3 * np.cos(0.2 * X)) * np.exp(-0.5 * VY ** 2)
             * (np.exp(-0.5 * (VX - 2.4) ** 2) + np.exp(-0.5 * (VX + 2.4) ** 2)) / (4 * math.pi))
    return np.broadcast_to(f, shape)


@pytest.mark.parametrize("idcase,lx,vmax", [(30, 4 * math.pi, 6.0), (31, 10 * math.pi, 8.0)])
def test_extension_init_closed_form(orc, idcase, lx, vmax):
    c = oracle.make_config((14, 11, 18, 13), [0, 0, -vmax, -vmax], [lx, lx, vmax, vmax], idcase, 0.05, 1)
    h = orc.init_case(c)
    assert_pinned(orc.interior_of(h), ext_closed_form(c, idcase))
    rest = h.copy()
    orc.interior_of(rest)[...] = 0.0
    assert not np.any(rest)


def test_strong_landau_density_closed_form(orc):
    """rho = (1 + 0.5 cos(x/2) + 0.5 cos(y/2)) to trapezoid accuracy on v in
    +-6 (the Gaussian's tails beyond 6 sigma are 2e-9)."""
    c = oracle.make_config((12, 10, 40, 36), [0, 0, -6, -6], [4 * math.pi] * 2 + [6, 6], 30, 0.05, 1)
    rho = orc.field_rho(c, orc.init_case(c))
    x, y, _, _ = axes(c)
    ref = 1 + 0.5 * np.cos(0.5 * x)[None, :] + 0.5 * np.cos(0.5 * y)[:, None]
    np.testing.assert_allclose(rho, ref, rtol=1e-8)


def test_two_stream_density_closed_form(orc):
    """Each beam integrates to 2 pi, so rho = 1 + 1e-3 cos(0.2 x), independent
    of y; the beams at +-2.4 sit 5.6 sigma inside v = +-8."""
    c = oracle.make_config((12, 10, 64, 48), [0, 0, -8, -8], [10 * math.pi] * 2 + [8, 8], 31, 0.05, 1)
    rho = orc.field_rho(c, orc.init_case(c))
    x = axes(c)[0]
    np.testing.assert_allclose(rho, np.broadcast_to(1 + 1e-3 * np.cos(0.2 * x)[None, :], rho.shape), rtol=2e-8)
    assert np.ptp(rho, axis=0).max() < 1e-14     # no y dependence


def test_extension_cases_run_the_whole_step(orc):
    """vo_create accepts 30 and 31 (the reference's dispatch would abort)."""
    for idcase, lx, vmax in ((30, 4 * math.pi, 6.0), (31, 10 * math.pi, 8.0)):
        c = oracle.make_config((16, 12, 14, 10), [0, 0, -vmax, -vmax], [lx, lx, vmax, vmax], idcase, 0.05, 2)
        st = orc.State(c)
        st.step(2)
        assert np.all(np.isfinite(st.diags()["nrj"]))
        st.close()
