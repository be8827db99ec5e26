"""Pins the oracle's spline prefilter and halo extension (rows A4/A5).

Reference: Spline::computeCoeff (VMO/Spline.hpp:114-199) on lines extended by
exchangeHalo/boundary_condition (VMO/Communication.hpp:713-898).  The
reference ships no numeric test for these; they are pinned here by
mathematics the recursions must satisfy, not by retyping them:

* interpolation identity -- from cp[n] = g[n] + a cp[n-1] (n >= -1) and
  c[n] = b cp[n] + a c[n+1] (n <= N) with a = sqrt3-2, b = sqrt3(1-a^2) it
  follows exactly that c[n-1] + 4 c[n] + c[n+1] = 6 g[n] for n in [-1, N]
  (a wrong a, b, sqrt3, loop range or index breaks it);
* exact periodic B-spline -- away from the line ends the coefficients equal
  the circulant solution (numpy FFT), the gap decaying like |a|^distance;
* boundary equations -- the two initial values (cp[-2] = g[-3] + g[-2],
  c[N+1] = sqrt3 (g[N+2] + g[N+1] + sum_k a^k g[N+1-k])) recovered from the
  output and compared with the alpha-sums built from f by the test itself;
* a constant line in 50-digit arithmetic (mpmath), by direct summation of the
  closed forms of both recursions;
* tensor structure -- the 2-D/4-D routines equal the 1-D operator applied
  along each axis (built from unit vectors);
* the halo exchange -- the reference's own start-up self-test pattern
  (VMO/Init.cpp:279-322) for the periodic layers, and a product input for the
  alpha-sum layers of every multiplicity.
"""
import math

import mpmath
import numpy as np
import pytest

from conftest import sld_box

ALPHA = math.sqrt(3) - 2


def M_of(n, orc):
    """(n+4) x n matrix of the 1-D operator, column j = spline_line(e_j)."""
    M = np.zeros((n + 4, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        M[:, j] = orc.spline_line(e)
    return M


def ext_line(f):
    """g[-3..N+2] (index 0 = position -3) from the periodic copy + alpha-sums."""
    n = f.size
    g = np.empty(n + 6)
    for i in range(-2, n + 2):
        g[i + 3] = f[i % n]
    g[0] = sum(ALPHA ** j * f[n - 2 - j] for j in range(1, 9))
    g[n + 5] = sum(ALPHA ** j * f[1 + j] for j in range(1, 9))
    return g


@pytest.mark.parametrize("n", [10, 12, 16, 33, 64])
def test_interpolation_identity_every_row(orc, n):
    rng = np.random.default_rng(n)
    f = rng.standard_normal(n)
    c = orc.spline_line(f)            # positions -2..n+1
    g = ext_line(f)                   # positions -3..n+2
    for pos in range(-1, n + 1):
        k = pos + 2
        lhs = (c[k - 1] + 4 * c[k] + c[k + 1]) / 6
        assert abs(lhs - g[pos + 3]) < 5e-15 * np.max(np.abs(f)) * 8, pos


@pytest.mark.parametrize("n", [32, 64, 128])
def test_matches_exact_periodic_spline_away_from_ends(orc, n):
    rng = np.random.default_rng(7 + n)
    f = rng.standard_normal(n)
    c = orc.spline_line(f)
    theta = 2 * np.pi * np.arange(n) / n
    cper = np.real(np.fft.ifft(np.fft.fft(f) * 6.0 / (4.0 + 2.0 * np.cos(theta))))
    fmax = np.max(np.abs(f))
    for pos in range(-2, n + 2):
        d = min(pos + 3, n + 2 - pos)         # distance to the alpha-sum cells
        gap = abs(c[pos + 2] - cper[pos % n])
        # truncation after MMAX=8 terms: |a|^9 at the ends, decaying like |a|^d
        assert gap <= 40 * abs(ALPHA) ** (8 + d) * fmax + 1e-14 * fmax, (pos, gap)
    if n >= 64:  # |a|^30 < 1e-17: the middle of a long line is the circulant solution
        mid = slice(30, n - 30)
        assert np.max(np.abs(c[2:-2][mid] - cper[mid])) < 1e-14 * fmax


@pytest.mark.parametrize("n", [10, 17, 40])
def test_boundary_equations(orc, n):
    rng = np.random.default_rng(100 + n)
    f = rng.standard_normal(n)
    c = orc.spline_line(f)
    g = ext_line(f)
    a = ALPHA
    b = math.sqrt(3) * (1 - a * a)
    # causal start: cp[-2] = (c[-2] - a c[-1]) / b = g[-3] + g[-2]
    cp_m2 = (c[0] - a * c[1]) / b
    assert abs(cp_m2 - (g[0] + g[1])) < 1e-14 * np.max(np.abs(f)) * 4
    # anti-causal start
    fnend = g[n + 5] + g[n + 4] + sum(a ** k * g[n + 4 - k] for k in range(1, n + 4))
    assert abs(c[n + 3] - math.sqrt(3) * fnend) < 1e-14 * np.max(np.abs(f)) * 4


@pytest.mark.parametrize("n", [10, 16])
def test_constant_line_closed_form(orc, n):
    mpmath.mp.dps = 50
    a = mpmath.sqrt(3) - 2
    b = mpmath.sqrt(3) * (1 - a * a)
    s8 = mpmath.fsum(a ** j for j in range(1, 9))
    # cp_k (position k-2) = sum_{i<=k} a^i + a^k s8 ; fnend = s8 + sum_{k=0}^{n+3} a^k
    cp = [mpmath.fsum(a ** i for i in range(k + 1)) + a ** k * s8 for k in range(n + 4)]
    cend = mpmath.sqrt(3) * (s8 + mpmath.fsum(a ** k for k in range(n + 4)))
    ref = []
    for k in range(n + 4):
        if k == n + 3:
            ref.append(cend)
        else:
            ref.append(b * mpmath.fsum(a ** (m - k) * cp[m] for m in range(k, n + 3)) + a ** (n + 3 - k) * cend)
    got = orc.spline_line(np.ones(n))
    for k in range(n + 4):
        assert abs(got[k] - float(ref[k])) < 2e-15, k  # a few ulp of accumulated rounding


def test_two_dim_is_tensor_product(orc):
    n = (12, 14, 10, 11)
    c = orc.make_config(*sld_box(n))
    rng = np.random.default_rng(3)
    h = orc.halo_zeros(c)
    orc.interior_of(h)[...] = rng.standard_normal(orc.interior_shape(c))
    f = np.array(orc.interior_of(h))
    orc.exchange_halo(c, h)
    h2 = h.copy()
    orc.spline_xy(c, h)
    Mx, My = M_of(n[0], orc), M_of(n[1], orc)
    # expected on x,y in [-2,N+1], interior v planes
    exp = np.einsum("xi,yj,abji->abyx", Mx, My, f)
    got = h[3:-3, 3:-3, 1:-1, 1:-1]
    assert np.max(np.abs(got - exp)) < 1e-13 * np.max(np.abs(exp))
    # v halo planes carry the same operator applied to the v-extended data
    orc.spline_vxvy(c, h2)
    Mvx, Mvy = M_of(n[2], orc), M_of(n[3], orc)
    exp4 = np.einsum("xi,yj,ak,bl,lkji->bayx", Mx, My, Mvx, Mvy, f)
    orc.spline_xy(c, h2)  # order of the two 2-D passes does not matter in exact arithmetic
    got4 = orc.ext_view(h2)
    assert np.max(np.abs(got4 - exp4)) < 1e-13 * np.max(np.abs(exp4))


def test_halo_selftest_pattern(orc):
    """The reference's start-up check (VMO/Init.cpp:279-322)."""
    n = (12, 13, 10, 11)
    c = orc.make_config(*sld_box(n))
    h = orc.halo_zeros(c)
    iy, ix = np.meshgrid(np.arange(n[1]), np.arange(n[0]), indexing="ij")
    for ivy in range(n[3]):
        for ivx in range(n[2]):
            h[ivy + 3, ivx + 3, 3:-3, 3:-3] = ((ivy + 111 * ivx) * 111 + iy) * 111 + ix
    orc.exchange_halo(c, h)
    for ivy in range(-2, n[3] + 2):
        for ivx in range(-2, n[2] + 2):
            jvy, jvx = ivy % n[3], ivx % n[2]
            ref = (((jvy + 111.0 * jvx) * 111.0 + (np.arange(-2, n[1] + 2) % n[1])[:, None]) * 111.0
                   + (np.arange(-2, n[0] + 2) % n[0])[None, :])
            got = h[ivy + 3, ivx + 3, 1:-1, 1:-1]
            assert np.max(np.abs(got - ref)) < 0.1


def test_halo_alpha_sums_product_input(orc):
    """For f = a(x) b(y) c(vx) d(vy) every halo cell is the product of the 1-D
    extensions, whatever the number of alpha-sum dims (1..4)."""
    n = (12, 11, 10, 13)
    c = orc.make_config(*sld_box(n))
    rng = np.random.default_rng(11)
    parts = [rng.uniform(0.5, 1.5, n[d]) for d in range(4)]
    h = orc.halo_zeros(c)
    orc.interior_of(h)[...] = np.einsum("l,k,j,i->lkji", parts[3], parts[2], parts[1], parts[0])
    orc.exchange_halo(c, h)
    ext = [ext_line(p) for p in parts]
    exp = np.einsum("l,k,j,i->lkji", ext[3], ext[2], ext[1], ext[0])
    assert np.max(np.abs(h - exp)) < 1e-14 * np.max(np.abs(exp)) * 8
