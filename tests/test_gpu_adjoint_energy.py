"""Gradient of the integrated field energy through the C-ABI
(vlp_adjoint_field_energy), and the adjoint workload at its stated size.

J(f^n) = dt * sum_{k=1..K} W(f^{n+k}),  W = sum_{x,y} (Ex^2 + Ey^2) dx dy of
the field solved from f^{n+k} at the end of step k (the nrj diagnostic is
log sqrt W, VMO/Diags.cpp:16-48).  The forward step is held to the oracle at
1e-12 elsewhere (test_gpu_parity.py, test_gpu_headline.py); these checks pin
the gradient to that forward map:

* J returned by the adjoint equals dt sum 2 (nrjx^2 + nrjy^2) of a plain run;
* 4-point central differences of J along random and density-modulated
  directions agree with <dJ/df, d> (relative 1e-8 on small grids);
* the velocity moment (vweight) equals the moment of the full gradient;
* the workload: a 96^4 FP64 Landau state (box 4 pi, v in [-6, 6], dt = 0.02)
  plus a density perturbation f^0 += M(v) (P delta)(x, y), M the unit
  Maxwellian, delta on a 64 x 64 grid drawn from N(0, 1e-3) (seeded), P the
  trigonometric interpolation 64 -> 96 (trig_interp_matrix); J over 50 steps
  with a checkpoint every 10 steps (the other states recomputed); dJ/ddelta =
  P^T (sum_v M dJ/df) P checked against 4-point central differences on 16
  random directions (relative 1e-6).

As in test_gpu_adjoint.py, the velocity clamp of computeFeet
(VMO/Advection.hpp:214-217) is not differentiable on the first velocity
plane; f there is e^{-18} of its peak on these boxes, far below the bars.
"""
import math

import numpy as np
import pytest

from conftest import sld_box

pytestmark = pytest.mark.gpu


def _energy_from_diags(sim, k0, K, dt):
    d = sim.diags()
    a, b = d["nrjx"][k0 + 1:k0 + K + 1], d["nrjy"][k0 + 1:k0 + K + 1]
    return dt * float(np.sum(2 * (a * a + b * b)))


@pytest.mark.parametrize("n,idcase,vmax,lx", [((16, 12, 20, 14), 10, 6.0, 4 * math.pi),
                                              ((12, 16, 14, 18), 20, 8.0, 10 * math.pi)])
def test_field_energy_gradient_matches_finite_differences(vlp, n, idcase, vmax, lx):
    rng = np.random.default_rng(31 + idcase)
    dt = 0.5 * (lx / n[0]) / vmax
    K = 5
    with vlp.Vlp4d(vlp.VlpConfig.make(*sld_box(n, vmax=vmax, dt=dt, nbiter=200, idcase=idcase, lx=lx))) as sim:
        sim.step(1)
        f0 = sim.get_f().copy()
        it0 = sim.iter
        J, grad = sim.field_energy_gradient(K, checkpoint_every=2)
        assert sim.iter == it0
        np.testing.assert_array_equal(sim.get_f(), f0)

        def Jf(f):
            sim.set_f(f)
            k0 = sim.iter
            sim.step(K)
            return _energy_from_diags(sim, k0, K, dt)

        J_run = Jf(f0)
        print(f"idcase {idcase}: J adjoint {J:.15e}, J run {J_run:.15e}")
        assert abs(J - J_run) <= 1e-13 * abs(J_run)
        xs = np.arange(n[0]) * (lx / n[0])
        k = 2 * math.pi / lx
        fmax = float(np.max(np.abs(f0)))
        for name, d in (("random", rng.standard_normal(f0.shape)),
                        ("density x", f0 * np.cos(k * xs)[None, None, None, :])):
            d = d * (fmax / float(np.max(np.abs(d))))
            eps = 1e-4
            fd = (8 * (Jf(f0 + eps * d) - Jf(f0 - eps * d)) - (Jf(f0 + 2 * eps * d) - Jf(f0 - 2 * eps * d))) / (12 * eps)
            ad = float(np.vdot(grad, d))
            rel = abs(fd - ad) / max(abs(fd), 1e-3 * float(np.sum(np.abs(grad * d))))
            print(f"idcase {idcase} {name}: finite difference {fd:.12e}, adjoint {ad:.12e}, rel {rel:.2e}")
            assert rel < 1e-8, name
        # the velocity moment against the full gradient
        w = rng.standard_normal((n[3], n[2]))
        sim.set_f(f0)
        J2, gm = sim.field_energy_gradient(K, vweight=w, checkpoint_every=2)
        assert abs(J2 - J) <= 1e-14 * abs(J)
        ref = np.einsum("ab,abyx->yx", w, grad)
        np.testing.assert_allclose(gm, ref, rtol=0, atol=1e-11 * np.max(np.abs(ref)))


def test_adjoint_workload_96_50_steps(vlp):
    """The adjoint workload at 96^4 over 50 steps (see the module docstring)."""
    n, lx, vmax, dt, K = 96, 4 * math.pi, 6.0, 0.02, 50
    ndir = 16
    rng = np.random.default_rng(20261203)
    delta = rng.normal(0.0, 1e-3, (64, 64))
    dirs = rng.standard_normal((ndir, 64, 64))
    P = vlp.trig_interp_matrix(64, n)
    v = -vmax + np.arange(n) * (2 * vmax / n)
    M = np.exp(-0.5 * (v[:, None] ** 2 + v[None, :] ** 2)) / (2 * math.pi)   # [vy][vx]
    nb = K * (4 * ndir + 2) + 1
    cfg = vlp.VlpConfig.make(*sld_box(n, vmax=vmax, dt=dt, nbiter=nb, idcase=10, lx=lx))
    with vlp.Vlp4d(cfg) as sim:
        base = sim.get_f()

        def f_of(dl):
            return base + M[:, :, None, None] * (P @ dl @ P.T)[None, None, :, :]

        def J_of(dl):
            sim.set_f(f_of(dl))
            k0 = sim.iter
            sim.step(K)
            return _energy_from_diags(sim, k0, K, dt)

        sim.set_f(f_of(delta))
        import time
        t0 = time.perf_counter()
        J, gm = sim.field_energy_gradient(K, vweight=M, checkpoint_every=10)
        t_adj = time.perf_counter() - t0
        gd = P.T @ gm @ P                                    # dJ/ddelta on the 64 x 64 grid
        J_run = J_of(delta)
        print(f"96^4 x{K}: J {J:.15e} (run {J_run:.15e}), adjoint sweep {t_adj:.2f} s", flush=True)
        assert abs(J - J_run) <= 1e-13 * abs(J_run)
        h = 1e-4
        worst = 0.0
        for r in range(ndir):
            d = dirs[r]
            fd = (8 * (J_of(delta + h * d) - J_of(delta - h * d)) - (J_of(delta + 2 * h * d) - J_of(delta - 2 * h * d))) / (
                12 * h)
            ad = float(np.vdot(gd, d))
            rel = abs(fd - ad) / abs(fd)
            worst = max(worst, rel)
            print(f"  direction {r:2d}: finite difference {fd:.12e}, adjoint {ad:.12e}, rel {rel:.2e}", flush=True)
        assert worst < 1e-6
