"""The adjoint of one step (vlp_adjoint_step) through the C-ABI.

The step maps f^n to f^{n+1}; vlp_adjoint_step returns (d f^{n+1}/d f^n)^T g.
The forward step it differentiates is itself held to the oracle at 1e-12
(test_gpu_parity.py), so the checks here pin the adjoint to the forward:

* idcase 2 (the reference's analytic field E = (-x, 0), VMO/Field.cpp:81-108):
  the feet do not depend on f, the step is linear, and the adjoint must be
  its exact transpose, <g, F d> = <F^T g, d> for any d, to round-off;
* idcases 10 and 20 (self-consistent field): the step is nonlinear through
  the feet; finite differences (4-point stencil) of L(f) = <g, F(f)> along
  random and density-modulated directions (which move the field, so the
  term through the Poisson solve and the half step is exercised) agree with
  <grad, d>;
* a torch.autograd.Function that binds the step and its adjoint passes
  torch.autograd.gradcheck (fast mode);
* the adjoint is reproducible: the coefficient adjoint is summed as 64-bit
  fixed-point integers and the E adjoint by gathers, so repeated calls (and
  the checkpointed sweep) return the same bits; the per-cell
  floating-point-atomic form (VLP4D_ADJ_BLOCK=0) agrees with it to
  round-off.

The velocity foot is clamped to the box, max(min(v - dt E, v_max), v_min)
(VMO/Advection.hpp:214-217), which has no derivative where v - dt E meets a
bound.  On the grid's first velocity plane (v = v_min) that happens wherever
the interpolated E vanishes, and the symmetric Landau states put cells there
to round-off; no other plane can switch branch unless |E| >= dx/dt =
2 v_max.  The adjoint takes the branch in force (a clamped foot contributes
no field term); the finite-difference checks give the cotangent zero weight
on the planes ivx = 0 and ivy = 0, where a difference quotient straddles
the corner.
* the adjoint leaves f and the step count alone, and the next step matches
  an untouched twin bit for bit.
"""
import math

import numpy as np
import pytest

from conftest import sld_box

pytestmark = pytest.mark.gpu


def _sim(vlp, n, idcase, vmax, lx, nbiter=200, precision=None):
    dt = 0.5 * (lx / n[0]) / vmax
    cfg = vlp.VlpConfig.make(*sld_box(n, vmax=vmax, dt=dt, nbiter=nbiter, idcase=idcase, lx=lx))
    return vlp.Vlp4d(cfg, precision=vlp.FP64 if precision is None else precision)


def _apply(sim, f):
    """One step from f; the context's step count advances."""
    sim.set_f(f)
    sim.step(1)
    return sim.get_f()


def _off_corner(g):
    """g with zero weight on the first vx and vy planes (array order vy, vx,
    y, x), where the velocity clamp can switch branch."""
    g = g.copy()
    g[0] = 0
    g[:, 0] = 0
    return g


def test_analytic_field_adjoint_is_exact_transpose(vlp):
    rng = np.random.default_rng(7)
    with _sim(vlp, (16, 12, 20, 14), 2, 6.0, 4 * math.pi) as sim:
        shape = sim.get_f().shape
        for _ in range(3):
            g = rng.standard_normal(shape)
            d = rng.standard_normal(shape)
            lhs = float(np.vdot(g, _apply(sim, d)))
            grad = sim.adjoint_step(g)
            rhs = float(np.vdot(grad, d))
            scale = float(np.sum(np.abs(g))) * float(np.max(np.abs(d)))
            print(f"idcase 2: <g, F d> = {lhs:.15e}, <F^T g, d> = {rhs:.15e}")
            assert abs(lhs - rhs) <= 1e-13 * scale


# (grid, idcase, vmax, box length)
NONLINEAR = [
    ((16, 12, 20, 14), 10, 6.0, 4 * math.pi),
    ((12, 16, 14, 18), 20, 8.0, 10 * math.pi),
]


@pytest.mark.parametrize("n,idcase,vmax,lx", NONLINEAR)
def test_adjoint_matches_central_differences(vlp, n, idcase, vmax, lx):
    rng = np.random.default_rng(11 + idcase)
    with _sim(vlp, n, idcase, vmax, lx) as sim:
        sim.step(2)  # a state with a developed field
        f0 = sim.get_f().copy()
        shape = f0.shape
        xs = np.arange(n[0]) * (lx / n[0])
        ys = np.arange(n[1]) * (lx / n[1])
        k = 2 * math.pi / lx
        dens = (np.cos(k * xs)[None, None, None, :] + 0.5 * np.sin(k * ys)[None, None, :, None])
        dirs = {
            "random": rng.standard_normal(shape),
            "density x": f0 * np.cos(k * xs)[None, None, None, :],
            "density xy": f0 * dens,
        }
        g = _off_corner(rng.standard_normal(shape))
        grad = sim.adjoint_step(g)
        np.testing.assert_array_equal(sim.get_f(), f0)
        fmax = float(np.max(np.abs(f0)))

        def L(f):
            return float(np.vdot(g, _apply(sim, f)))

        for name, d in dirs.items():
            d = d * (fmax / float(np.max(np.abs(d))))
            eps = 1e-4
            # 4-point stencil: truncation O(eps^4)
            fd = (8 * (L(f0 + eps * d) - L(f0 - eps * d)) - (L(f0 + 2 * eps * d) - L(f0 - 2 * eps * d))) / (12 * eps)
            ad = float(np.vdot(grad, d))
            # against the sum of the terms' sizes too, should <grad, d> cancel
            den = max(abs(fd), 1e-2 * float(np.sum(np.abs(grad * d))))
            rel = abs(fd - ad) / den
            print(f"idcase {idcase} {name}: finite difference {fd:.12e}, adjoint {ad:.12e}, rel {rel:.2e}")
            assert rel < 1e-9, name


def test_gradcheck_through_autograd_function(vlp):
    torch = pytest.importorskip("torch")
    n, idcase, vmax, lx = NONLINEAR[0]
    with _sim(vlp, n, idcase, vmax, lx, nbiter=400) as sim:
        sim.step(1)
        f0 = sim.get_f().copy()

        class Step(torch.autograd.Function):
            @staticmethod
            def forward(ctx, f):
                ctx.save_for_backward(f)
                return torch.from_numpy(_apply(sim, f.detach().numpy()).copy())

            @staticmethod
            def backward(ctx, gout):
                (f,) = ctx.saved_tensors
                sim.set_f(f.detach().numpy())
                return torch.from_numpy(sim.adjoint_step(gout.detach().numpy()).copy())

        mask = torch.from_numpy(_off_corner(np.ones_like(f0)))
        f = torch.tensor(f0, dtype=torch.float64, requires_grad=True)
        assert torch.autograd.gradcheck(lambda x: Step.apply(x) * mask, (f,), eps=1e-6, atol=1e-9, rtol=1e-7,
                                        nondet_tol=1e-12, fast_mode=True)


def test_adjoint_leaves_the_run_unchanged(vlp):
    n, idcase, vmax, lx = NONLINEAR[1]
    g = np.random.default_rng(3).standard_normal(n[::-1])
    with _sim(vlp, n, idcase, vmax, lx, nbiter=6) as a, _sim(vlp, n, idcase, vmax, lx, nbiter=6) as b:
        a.step(2)
        b.step(2)
        a.adjoint_step(g)
        assert a.iter == 2
        np.testing.assert_array_equal(a.get_f(), b.get_f())
        a.step(2)
        b.step(2)
        np.testing.assert_array_equal(a.get_f(), b.get_f())
        da, db = a.diags(), b.diags()
        for key in ("nrj", "nrjx", "nrjy", "mass", "l2norm"):
            np.testing.assert_array_equal(da[key], db[key])


@pytest.mark.parametrize("k", [0, 1])
def test_adjoint_is_reproducible(vlp, monkeypatch, k):
    n, idcase, vmax, lx = NONLINEAR[k]
    rng = np.random.default_rng(11 + k)
    g = rng.standard_normal(n[::-1])
    with _sim(vlp, n, idcase, vmax, lx, nbiter=8) as sim:
        sim.step(2)
        a = sim.adjoint_step(g)
        b = sim.adjoint_step(g)
        np.testing.assert_array_equal(a, b)
        s3a = sim.adjoint_steps(3, g)
        s3b = sim.adjoint_steps(3, g)
        np.testing.assert_array_equal(s3a, s3b)
        # the floating-point-atomic form gives the same gradient to round-off
        monkeypatch.setenv("VLP4D_ADJ_BLOCK", "0")
        c = sim.adjoint_step(g)
        monkeypatch.delenv("VLP4D_ADJ_BLOCK")
        np.testing.assert_allclose(c, a, rtol=0, atol=1e-12 * np.abs(a).max())


def test_adjoint_rejects_bad_requests(vlp):
    n, idcase, vmax, lx = NONLINEAR[0]
    with _sim(vlp, n, idcase, vmax, lx, precision=vlp.FP32) as sim:
        with pytest.raises(vlp.Vlp4dError):
            sim.adjoint_step(np.zeros(n[::-1]))
    with _sim(vlp, n, idcase, vmax, lx) as sim:
        with pytest.raises(vlp.Vlp4dError):
            sim.adjoint_step(np.zeros(10))


def test_checkpointed_sweep_matches_chained_steps(vlp):
    # vlp_adjoint_steps(K) against K one-step adjoints chained by hand from
    # stored states (a composition independent of the checkpoint schedule);
    # K = 5 has segments of 3 and 2
    n, idcase, vmax, lx = NONLINEAR[1]
    rng = np.random.default_rng(17)
    with _sim(vlp, n, idcase, vmax, lx, nbiter=20) as sim:
        sim.step(1)
        f0 = sim.get_f().copy()
        g = _off_corner(rng.standard_normal(f0.shape))
        np.testing.assert_allclose(sim.adjoint_steps(1, g), sim.adjoint_step(g), rtol=0, atol=1e-12 * np.abs(g).max())
        K = 5
        states = [f0]
        for _ in range(K - 1):
            states.append(_apply(sim, states[-1]))
        sim.set_f(f0)
        lam = g
        for k in range(K - 1, -1, -1):
            sim.set_f(states[k])
            lam = sim.adjoint_step(lam)
        sim.set_f(f0)
        it = sim.iter
        got = sim.adjoint_steps(K, g)
        assert sim.iter == it
        np.testing.assert_array_equal(sim.get_f(), f0)
        rel = np.max(np.abs(got - lam)) / np.max(np.abs(lam))
        print(f"checkpointed {K}-step adjoint vs chained: rel {rel:.2e}")
        assert rel < 1e-12


def test_checkpointed_sweep_matches_finite_differences(vlp):
    n, idcase, vmax, lx = NONLINEAR[1]
    rng = np.random.default_rng(23)
    K = 3
    with _sim(vlp, n, idcase, vmax, lx, nbiter=60) as sim:
        sim.step(1)
        f0 = sim.get_f().copy()
        it0 = sim.iter
        g = _off_corner(rng.standard_normal(f0.shape))
        grad = sim.adjoint_steps(K, g)
        d = rng.standard_normal(f0.shape)
        d *= np.max(np.abs(f0)) / np.max(np.abs(d))

        def L(f):
            sim.set_f(f)
            sim.step(K)
            v = float(np.vdot(g, sim.get_f()))
            return v

        eps = 1e-4
        fd = (8 * (L(f0 + eps * d) - L(f0 - eps * d)) - (L(f0 + 2 * eps * d) - L(f0 - 2 * eps * d))) / (12 * eps)
        ad = float(np.vdot(grad, d))
        rel = abs(fd - ad) / abs(fd)
        print(f"{K}-step adjoint: finite difference {fd:.12e}, adjoint {ad:.12e}, rel {rel:.2e}")
        assert rel < 1e-8
        assert it0 < sim.iter
