"""Pins the oracle's density, Poisson solve and field diagnostics (A7-A10).

Reference: field_rho/field_poisson (VMO/Field.cpp:5-137), Efield
(VMO/Efield.cpp:3-106), Impl::FFT (lib/OpenMP_FFT.hpp:106-389), Diags::compute
(VMO/Diags.cpp:16-48).  The reference's FFT is FFTW3 (a system package with
no pinned version, absent here); the oracle's DFT is pinned against numpy's
pocketfft, an independent FFT with the same conventions (unnormalised r2c
along x then c2c along y; c2r discards the imaginary part of the DC and
Nyquist bins).  FFTW's own last-bit rounding stays unpinned.  The Poisson
solve and the diagnostics are pinned by closed forms.
"""
import math

import numpy as np
import pytest

from conftest import sld_box


@pytest.mark.parametrize("shape", [(16, 16), (12, 20), (15, 9)])
def test_rfft2_matches_numpy(orc, shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape)  # (ny, nx)
    got = orc.rfft2(x)
    ref = np.fft.rfft2(x)
    assert np.max(np.abs(got - ref)) < 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("shape", [(16, 16), (12, 20), (15, 9)])
def test_irfft2_matches_numpy_including_dc_nyquist_rule(orc, shape):
    ny, nx = shape
    rng = np.random.default_rng(3 * nx + ny)
    X = rng.standard_normal((ny, nx // 2 + 1)) + 1j * rng.standard_normal((ny, nx // 2 + 1))
    got = orc.irfft2(X, nx)
    ref = np.fft.irfft2(X, s=(ny, nx)) * nx * ny  # the reference's transforms are unnormalised
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))


def _poisson(orc, n, lx, ly, rho):
    c = orc.make_config([n[0], n[1], 16, 16], [0, 0, -6, -6], [lx, ly, 6, 6], 10, 0.01, 1)
    return c, orc.field_poisson(c, rho)


@pytest.mark.parametrize("mx,my,amp", [(1, 0, 0.05), (0, 2, -0.3), (2, 1, 0.1), (3, -2, 0.2)])
def test_poisson_single_mode(orc, mx, my, amp):
    # rho = 1 + A cos(kx x + ky y): E = -grad phi, -lap phi = rho - 1
    n, lx, ly = (32, 24), 4 * math.pi, 6 * math.pi
    x = np.arange(n[0]) * lx / n[0]
    y = np.arange(n[1]) * ly / n[1]
    kx, ky = 2 * math.pi * mx / lx, 2 * math.pi * my / ly
    ph = kx * x[None, :] + ky * y[:, None]
    rho = 1 + amp * np.cos(ph)
    c, (phi, ex, ey) = _poisson(orc, n, lx, ly, rho)
    k2 = kx * kx + ky * ky
    assert np.max(np.abs(phi - amp * np.cos(ph) / k2)) < 1e-13
    assert np.max(np.abs(ex - amp * kx * np.sin(ph) / k2)) < 1e-13
    assert np.max(np.abs(ey - amp * ky * np.sin(ph) / k2)) < 1e-13


def test_poisson_nyquist_mode_in_x(orc):
    # rho - 1 = A (-1)^ix: phi keeps the mode, Ex loses it (imaginary Nyquist bin dropped by c2r)
    n, lx, ly = (16, 12), 4 * math.pi, 4 * math.pi
    amp = 0.02
    rho = 1 + amp * ((-1.0) ** np.arange(n[0]))[None, :] * np.ones((n[1], 1))
    c, (phi, ex, ey) = _poisson(orc, n, lx, ly, rho)
    kn = (n[0] // 2) * 2 * math.pi / lx
    assert np.max(np.abs(phi - (rho - 1) / kn ** 2)) < 1e-14
    assert np.max(np.abs(ex)) < 1e-15 and np.max(np.abs(ey)) < 1e-15


def test_analytic_field_case2_keeps_rho(orc):
    n = (12, 10)
    c = orc.make_config([n[0], n[1], 12, 12], [-1, -1, -1, -1], [1, 1, 1, 1], 2, 0.01, 1)
    rho = np.full((n[1], n[0]), 0.7)
    phi, ex, ey = orc.field_poisson(c, rho)
    x = -1 + np.arange(n[0]) * (2 / n[0])
    assert np.array_equal(phi, rho)
    assert np.max(np.abs(ex + x[None, :])) == 0 and np.max(np.abs(ey)) == 0


def test_diags_closed_form(orc):
    # E = (A sin(kx x), B cos(ky y)) on full periods: sum E^2 dxdy = (A^2 + B^2) Lx Ly / 2
    n, lx, ly = (20, 16), 4 * math.pi, 2 * math.pi
    c = orc.make_config([n[0], n[1], 12, 12], [0, 0, -6, -6], [lx, ly, 6, 6], 10, 0.01, 1)
    x = np.arange(n[0]) * lx / n[0]
    y = np.arange(n[1]) * ly / n[1]
    A, B = 0.3, 0.2
    ex = A * np.sin(2 * math.pi * x / lx)[None, :] * np.ones((n[1], 1))
    ey = B * np.cos(2 * math.pi * 2 * y / ly)[:, None] * np.ones((1, n[0]))
    phi = 0.1 * np.cos(2 * math.pi * x / lx)[None, :] * np.ones((n[1], 1))
    nrj, nrjx, nrjy, mass = orc.diags_field(c, phi, ex, ey)
    area = lx * ly
    assert abs(nrj - math.log(math.sqrt((A * A + B * B) * area / 2))) < 1e-14
    assert abs(nrjx - math.sqrt(0.5 * A * A * area / 2)) < 1e-14
    assert abs(nrjy - math.sqrt(0.5 * B * B * area / 2)) < 1e-14
    assert abs(mass) < 1e-14
    # zero field: the reference reports -1e9 instead of log(0) (VMO/Diags.cpp:43)
    z = np.zeros_like(ex)
    assert orc.diags_field(c, phi, z, z)[0] == -1e9


def test_rho_of_sld10_initial_state(orc):
    # sum_v exp(-v^2/2)/(2 pi) dv^2 on v in [-9, 9) with dv = 18/N is 1 to
    # round-off (Gaussian tails e^-40 and aliasing e^{-2 pi^2/dv^2}), so
    # rho = 1 + 0.05 cos(x/2) cos(y/2) at the grid points (VMO/Init.cpp:226-245).
    n = (16, 12, 64, 64)
    c = orc.make_config(*sld_box(n, vmax=9.0))
    h = orc.init_case(c)
    rho = orc.field_rho(c, h)
    x = np.arange(n[0]) * c.dx[0]
    y = np.arange(n[1]) * c.dx[1]
    ref = 1 + 0.05 * np.cos(0.5 * x)[None, :] * np.cos(0.5 * y)[:, None]
    assert np.max(np.abs(rho - ref)) < 2e-14


def test_rho_of_tsi20_initial_state(orc):
    # two beams at +-2.4 (vthx 1) times a vy Gaussian (vthy 0.5): density
    # 1 - 0.05 sin(0.2 x) - 0.25 sin(0.2 y) (VMO/Init.cpp:247-277)
    n = (12, 10, 96, 128)
    lx = 10 * math.pi
    c = orc.make_config(n, [0, 0, -9, -9], [lx, lx, 9, 9], 20, 0.01, 1)
    h = orc.init_case(c)
    rho = orc.field_rho(c, h)
    x = np.arange(n[0]) * c.dx[0]
    y = np.arange(n[1]) * c.dx[1]
    ref = 1 - 0.05 * np.sin(0.2 * x)[None, :] - 0.25 * np.sin(0.2 * y)[:, None]
    # the beam at -2.4 is cut at v = -9: Gaussian tail mass ~ e^{-6.6^2/2} ~ 3e-10
    assert np.max(np.abs(rho - ref)) < 1e-9


def test_l2norm_closed_form(orc):
    n = (10, 11, 12, 13)
    c = orc.make_config(*sld_box(n))
    h = orc.halo_zeros(c)
    orc.interior_of(h)[...] = 2.0
    vol = np.prod([c.nxmax[d] * c.dx[d] for d in range(4)])
    assert abs(orc.l2norm(c, h) - math.sqrt(4.0 * vol)) < 1e-12
