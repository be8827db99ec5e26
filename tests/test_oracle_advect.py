"""Pins the oracle's advection steps and characteristic feet (A6, A11, A12).

Reference: interp_2D/advect_2D_xy (VMO/Advection.hpp:87-164, 299-406),
computeFeet (:166-218), interp_4D/advect_4D (:221-296, 409-498).

The reference writes the cubic B-spline weights as eta0..eta3 expressions and
the E interpolation as hand-expanded Lagrange polynomials.  The checks here
evaluate the same interpolants from their textbook definitions instead:
* the centred cubic B-spline B3(t) summed over the stencil, sum_k c[k] B3(x-k);
* Lagrange basis products prod_{m!=j} (s-m)/(j-m) on the three E nodes;
plus closed forms: dt = 0 reproduces f through the node identity, a uniform
E gives x* = x - dt (v - dt E/2), v* = v - dt E.
"""
import math

import numpy as np
import pytest

from conftest import sld_box


def B3(t):
    t = np.abs(t)
    return np.where(t < 1, 2.0 / 3.0 - t * t + 0.5 * t ** 3, np.where(t < 2, (2 - t) ** 3 / 6.0, 0.0))


def lagrange3(s):
    """weights of nodes 0,1,2 at s (textbook product form)."""
    nodes = [0.0, 1.0, 2.0]
    out = []
    for j in range(3):
        w = 1.0
        for m in range(3):
            if m != j:
                w *= (s - nodes[m]) / (nodes[j] - nodes[m])
        out.append(w)
    return out


def _state(orc, n, seed, vmax=6.0, dt=0.05):
    c = orc.make_config(*sld_box(n, vmax=vmax, dt=dt))
    rng = np.random.default_rng(seed)
    h = orc.init_case(c)
    orc.interior_of(h)[...] *= 1 + 0.2 * rng.standard_normal(orc.interior_shape(c))
    return c, h


def test_advect2d_is_bspline_evaluation(orc):
    n = (14, 12, 10, 11)
    c, h = _state(orc, n, 5)
    orc.exchange_halo(c, h)
    orc.spline_xy(c, h)
    cxy = h.copy()
    dt = 0.5 * c.dt
    assert orc.advect_2d(c, h, dt) == 0
    got = orc.interior_of(h)
    x = c.minPhy[0] + np.arange(n[0]) * c.dx[0]
    y = c.minPhy[1] + np.arange(n[1]) * c.dx[1]
    for ivy in range(n[3]):
        vy = c.minPhy[3] + ivy * c.dx[3]
        for ivx in range(n[2]):
            vx = c.minPhy[2] + ivx * c.dx[2]
            fx = (x - dt * vx - c.minPhy[0]) / c.dx[0]
            fy = (y - dt * vy - c.minPhy[1]) / c.dx[1]
            plane = cxy[ivy + 3, ivx + 3]      # y, x halo plane: position p at index p+3
            kx = np.arange(-2, n[0] + 2)
            ky = np.arange(-2, n[1] + 2)
            Wx = B3(fx[:, None] - kx[None, :])   # (n0, n0+4)
            Wy = B3(fy[:, None] - ky[None, :])
            ref = Wy @ plane[1:-1, 1:-1] @ Wx.T
            assert np.max(np.abs(got[ivy, ivx] - ref)) < 2e-15 * np.max(np.abs(plane)) * 8


def test_advect2d_reports_too_large_step(orc):
    n = (12, 12, 10, 10)
    c, h = _state(orc, n, 1)
    orc.exchange_halo(c, h)
    orc.spline_xy(c, h)
    # half-step displacement of 3 cells: the reference exits "Time step is too large"
    assert orc.advect_2d(c, h, 3 * c.dx[0] / c.maxPhy[2] * 2) == orc.VO_E_CFL


def test_feet_match_textbook_lagrange(orc):
    n = (16, 12, 10, 14)
    c = orc.make_config(*sld_box(n, dt=0.07))
    rng = np.random.default_rng(9)
    ex = rng.standard_normal((n[1], n[0]))
    ey = rng.standard_normal((n[1], n[0]))
    dt = c.dt
    for (ix, iy, ivx, ivy) in [(0, 0, 0, 0), (15, 11, 9, 13), (3, 7, 4, 2), (8, 1, 6, 9), (12, 5, 0, 13)]:
        x = c.minPhy[0] + ix * c.dx[0]
        y = c.minPhy[1] + iy * c.dx[1]
        vx = c.minPhy[2] + ivx * c.dx[2]
        vy = c.minPhy[3] + ivy * c.dx[3]
        sx = (x - 0.5 * dt * vx - c.minPhy[0]) / c.dx[0]
        sy = (y - 0.5 * dt * vy - c.minPhy[1]) / c.dx[1]
        px, py = int(round(sx)) - 1, int(round(sy)) - 1
        wx, wy = lagrange3(sx - px), lagrange3(sy - py)
        Ex = sum(wx[a] * wy[b] * ex[(py + b) % n[1], (px + a) % n[0]] for a in range(3) for b in range(3))
        Ey = sum(wx[a] * wy[b] * ey[(py + b) % n[1], (px + a) % n[0]] for a in range(3) for b in range(3))
        ref = [x - dt * (vx - 0.5 * dt * Ex), y - dt * (vy - 0.5 * dt * Ey),
               min(max(vx - dt * Ex, c.minPhy[2]), c.maxPhy[2]), min(max(vy - dt * Ey, c.minPhy[3]), c.maxPhy[3])]
        got = orc.feet(c, ex, ey, ix, iy, ivx, ivy, dt)
        assert np.max(np.abs(got - np.array(ref))) < 1e-14 * 16


def test_feet_uniform_field_closed_form(orc):
    n = (12, 12, 12, 12)
    c = orc.make_config(*sld_box(n, dt=0.05))
    E0 = (0.3, -0.2)
    ex = np.full((n[1], n[0]), E0[0])
    ey = np.full((n[1], n[0]), E0[1])
    for idx in [(0, 0, 5, 5), (11, 6, 3, 8)]:
        ix, iy, ivx, ivy = idx
        x, y = ix * c.dx[0], iy * c.dx[1]
        vx, vy = c.minPhy[2] + ivx * c.dx[2], c.minPhy[3] + ivy * c.dx[3]
        got = orc.feet(c, ex, ey, *idx, c.dt)
        ref = [x - c.dt * (vx - 0.5 * c.dt * E0[0]), y - c.dt * (vy - 0.5 * c.dt * E0[1]),
               vx - c.dt * E0[0], vy - c.dt * E0[1]]
        assert np.max(np.abs(got - np.array(ref))) < 1e-14


@pytest.mark.parametrize("seed", [1, 2])
def test_advect4d_is_tensor_bspline_at_feet(orc, seed):
    n = (12, 10, 11, 12)
    c, h = _state(orc, n, seed, dt=0.06)
    rng = np.random.default_rng(seed + 10)
    ex = 0.3 * rng.standard_normal((n[1], n[0]))
    ey = 0.3 * rng.standard_normal((n[1], n[0]))
    orc.exchange_halo(c, h)
    orc.spline_xy(c, h)
    orc.spline_vxvy(c, h)
    c4 = h.copy()
    assert orc.advect_4d(c, ex, ey, h, c.dt) == 0
    got = orc.interior_of(h)
    for (ix, iy, ivx, ivy) in [(0, 0, 0, 0), (11, 9, 10, 11), (5, 4, 0, 11), (3, 8, 5, 6), (7, 2, 9, 1)]:
        xs = orc.feet(c, ex, ey, ix, iy, ivx, ivy, c.dt)
        f = [(xs[d] - c.minPhy[d]) / c.dx[d] for d in range(4)]
        ks = [np.arange(int(math.floor(f[d])) - 1, int(math.floor(f[d])) + 3) for d in range(4)]
        w = [B3(f[d] - ks[d]) for d in range(4)]
        blk = c4[np.ix_(ks[3] + 3, ks[2] + 3, ks[1] + 3, ks[0] + 3)]
        ref = np.einsum("l,k,j,i,lkji->", w[3], w[2], w[1], w[0], blk)
        assert abs(got[ivy, ivx, iy, ix] - ref) < 1e-14 * np.max(np.abs(blk)) * 16


def test_advect4d_zero_step_reproduces_interior(orc):
    # dt = 0: every foot is its node, weights (1/6, 2/3, 1/6, 0) in all dims,
    # and c[n-1] + 4 c[n] + c[n+1] = 6 f[n] holds on every interior node
    # (test_oracle_spline), so the step returns f everywhere.
    n = (24, 24, 24, 24)
    c, h = _state(orc, n, 3, dt=0.0)
    f0 = np.array(orc.interior_of(h))
    zero = np.zeros((n[1], n[0]))
    orc.exchange_halo(c, h)
    orc.spline_xy(c, h)
    orc.spline_vxvy(c, h)
    assert orc.advect_4d(c, zero, zero, h, 0.0) == 0
    got = orc.interior_of(h)
    assert np.max(np.abs(got - f0)) < 1e-13 * np.max(np.abs(f0))


def test_advect4d_reports_foot_outside(orc):
    n = (12, 12, 10, 10)
    c, h = _state(orc, n, 4, dt=0.05)
    zero = np.zeros((n[1], n[0]))
    orc.exchange_halo(c, h)
    orc.spline_xy(c, h)
    orc.spline_vxvy(c, h)
    assert orc.advect_4d(c, zero, zero, h, 3 * c.dx[0] / c.maxPhy[2]) == orc.VO_E_CFL
