"""The Lagrange split-step oracle (oracle/vlp4d_lag.c) pinned by what the
mathematics fixes, independently of its own formulas:

* the order-5 basis of VO/advect.hpp:72-86 is the identity at the nodes
  (px = 2 and px = 3, which also makes the interpolant continuous from one
  cell to the next), a partition of unity, and reproduces every polynomial of
  degree <= 4 exactly (degree 5 it does not: a property of the reference's
  formula, checked too, so that a retyped "textbook" Lagrange basis fails);
* a 1-D advection whose displacement is a whole number of cells is a
  periodic roll of the data (fixes the sign of the foot, the fmod wrap and
  the stencil offset), along x/y (dt*v) and along vx/vy (dt*E);
* a smooth periodic profile advected by a fraction of a cell matches the
  translated profile to the scheme's order;
* every 1-D advection moves whole lines by one shift, so with a partition of
  unity the total of f is conserved to round-off by a full step;
* rho is dvx*dvy times the velocity sum;
* the oracle's onetimestep is the composition x/2, y/2, rho + Poisson, vx, vy, x/2,
  y/2 of VO/timestep.hpp:11-55, with the diagnostics of the mid-step field.
The reference's own tests for this miniapp hold no golden vectors.
"""
import math

import numpy as np
import pytest

import oracle


@pytest.fixture(scope="module", autouse=True)
def _lib():
    oracle.load()


def _cfg(n=(16, 12, 10, 14), lx=4 * math.pi, ly=None, vmin=-6.0, vmax=6.0, idcase=10, dt=0.05, nbiter=8):
    ly = lx if ly is None else ly
    return oracle.make_config(list(n), [0.0, 0.0, vmin, vmin], [lx, ly, vmax, vmax], idcase, dt, nbiter)


def test_basis_nodes_partition_and_polynomials():
    k = np.arange(6.0)
    np.testing.assert_allclose(oracle.lag_basis(2.0), np.eye(6)[2], atol=1e-15)
    np.testing.assert_allclose(oracle.lag_basis(3.0), np.eye(6)[3], atol=1e-15)
    for px in np.linspace(2.0, 3.0, 37):
        c = oracle.lag_basis(px)
        assert abs(c.sum() - 1.0) < 1e-14
        for m in range(5):
            assert abs(np.dot(c, k ** m) - px ** m) < 1e-12 * max(1.0, px ** m), (px, m)
    # degree 5 is not reproduced by the reference's basis
    # (the midpoint is exact by symmetry; measured max 0.42 over [2, 3])
    assert max(abs(np.dot(oracle.lag_basis(p), k ** 5) - p ** 5) for p in np.linspace(2.0, 3.0, 41)) > 0.1


@pytest.mark.parametrize("d", [0, 1])
def test_whole_cell_shift_is_a_roll_in_space(d):
    # v = -6..5 (dv = 1), dt = dx: the foot is v cells upstream
    n = (16, 12, 12, 12)
    c = _cfg(n=n, lx=16.0, ly=12.0, vmin=-6.0, vmax=6.0)
    f = np.random.default_rng(d).standard_normal(n[::-1])
    dx = c.dx[d]
    out = oracle.lag_advect_xy(c, d, f, dx)
    for iv in range(n[2 + d]):
        v = int(round(c.minPhy[2 + d] + iv * c.dx[2 + d]))
        # array axes (vy, vx, y, x): x advection by vx slices (vy, y, x),
        # y advection by vy slices (vx, y, x)
        sl, axis = ((slice(None), iv), 2) if d == 0 else ((iv,), 1)
        np.testing.assert_array_equal(out[sl], np.roll(f[sl], v, axis=axis))


@pytest.mark.parametrize("d", [2, 3])
def test_whole_cell_shift_is_a_roll_in_velocity(d):
    n = (8, 6, 16, 14)
    c = _cfg(n=n, vmin=-4.0, vmax=4.0)
    f = np.random.default_rng(d).standard_normal(n[::-1])
    dv, dt = c.dx[d], 0.1
    shifts = np.random.default_rng(7 + d).integers(-3, 4, size=(n[1], n[0]))
    e = shifts * dv / dt
    out = oracle.lag_advect_v(c, d, f, e, dt)
    axis = 1 if d == 2 else 0
    for iy in range(n[1]):
        for ix in range(n[0]):
            col = f[:, :, iy, ix]
            np.testing.assert_allclose(out[:, :, iy, ix], np.roll(col, int(shifts[iy, ix]), axis=axis), atol=1e-13)


def test_fractional_shift_translates_a_smooth_profile():
    n = (64, 10, 10, 10)
    lx = 2 * math.pi
    c = _cfg(n=n, lx=lx, vmin=-1.0, vmax=1.0)
    x = np.arange(n[0]) * c.dx[0]
    f = np.broadcast_to(np.sin(x) + 0.5 * np.cos(3 * x), n[::-1]).copy()
    dt = 0.37 * c.dx[0]
    out = oracle.lag_advect_xy(c, 0, f, dt)
    for iv in range(n[2]):
        v = c.minPhy[2] + iv * c.dx[2]
        xs = x - v * dt
        exact = np.sin(xs) + 0.5 * np.cos(3 * xs)
        # measured 2.9e-6 (O(dx^5)); a reversed foot would miss by 0.1
        assert np.max(np.abs(out[0, iv, 0] - exact)) < 1e-5


def test_field_rho_is_the_velocity_sum():
    c = _cfg()
    f = np.random.default_rng(3).standard_normal(tuple(c.nxmax[d] for d in (3, 2, 1, 0)))
    rho = oracle.lag_field_rho(c, f)
    np.testing.assert_allclose(rho, f.sum(axis=(0, 1)) * c.dx[2] * c.dx[3], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("idcase", [10, 20])
def test_step_conserves_the_total_and_composes_the_phases(idcase):
    lx = 4 * math.pi if idcase == 10 else 10 * math.pi
    c = _cfg(n=(16, 12, 20, 14), lx=lx, vmin=-6.0, vmax=6.0, idcase=idcase, dt=0.1, nbiter=3)
    st = oracle.LagState(c)
    f0 = st.f()
    np.testing.assert_array_equal(f0, oracle.lag_init_case(c))
    st.step(1)
    f1 = st.f()
    assert abs(f1.sum() - f0.sum()) < 1e-13 * np.abs(f0).sum()
    # the same step from the phase functions
    dt = c.dt
    g = oracle.lag_advect_xy(c, 1, oracle.lag_advect_xy(c, 0, f0, 0.5 * dt), 0.5 * dt)
    phi, ex, ey = oracle.field_poisson(c, oracle.lag_field_rho(c, g))
    g = oracle.lag_advect_v(c, 3, oracle.lag_advect_v(c, 2, g, ex, dt), ey, dt)
    g = oracle.lag_advect_xy(c, 1, oracle.lag_advect_xy(c, 0, g, 0.5 * dt), 0.5 * dt)
    np.testing.assert_array_equal(f1, g)
    d = st.diags()
    ref = oracle.diags_field(c, phi, ex, ey)
    assert d["nrj"][1] == ref[0] and d["mass"][1] == ref[3]
    np.testing.assert_array_equal(st.field(1), ex)
    st.close()
