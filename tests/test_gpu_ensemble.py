"""Ensembles (one captured HIP graph per two steps of every member) and the
4-D advection autotuner, through the C-ABI.

The ensemble runs each member's kernels, in each member's order, from a
replayed graph; every replay is checked against a twin set of contexts
stepped eagerly with vlp_step, with eager work on the members between the
replays (an odd step count, which runs its last step eagerly and leaves the
members' buffers swapped, so the next replay needs a fresh capture).  The
autotuner changes the 4-D advection's workgroup shape, which does not
change any cell's arithmetic, and the spline passes' kernels (128- or
64-position LDS tiles or line kernels), which split the line recursions
differently: a tuned context gives the untuned one's result to round-off
(FP64) or float rounding (float storage).
"""
import math
import os

import numpy as np
import pytest

from conftest import relmax, sld_box

pytestmark = pytest.mark.gpu

# uneven grids and both storage precisions; CFL-valid dt
MEMBERS = [
    ((24, 20, 16, 18), 10, "FP64", 6.0, 4 * math.pi),
    ((32, 16, 24, 14), 20, "FP32", 9.0, 10 * math.pi),
    ((16, 12, 20, 14), 10, "FP64", 6.0, 4 * math.pi),
    ((32, 32, 32, 16), 10, "FP32_ACC", 6.0, 4 * math.pi),
]


def _make(vlp, nbiter):
    out = []
    for n, idcase, prec, vmax, lx in MEMBERS:
        dt = 0.5 * (lx / n[0]) / vmax
        cfg = vlp.VlpConfig.make(*sld_box(n, vmax=vmax, dt=dt, nbiter=nbiter, idcase=idcase, lx=lx))
        out.append(vlp.Vlp4d(cfg, precision=getattr(vlp, prec)))
    return out


def test_ensemble_replays_match_eager_steps(vlp):
    nbiter = 16
    ens_members = _make(vlp, nbiter)
    eager = _make(vlp, nbiter)
    try:
        with vlp.Ensemble(ens_members) as ens:
            assert len(ens) == len(MEMBERS)
            done = 0
            # 4 = two replays; 3 = one replay + an eager step (buffers swapped);
            # then an eager vlp_step on every member; then replays again
            for k, how in ((4, "ens"), (3, "ens"), (1, "eager"), (4, "ens"), (2, "ens")):
                if how == "ens":
                    ens.step(k)
                else:
                    for m in ens_members:
                        m.step(k)
                for m in eager:
                    m.step(k)
                done += k
                for a, b in zip(ens_members, eager):
                    assert a.iter == b.iter == done
                    fa, fb = a.get_f(), b.get_f()
                    assert np.array_equal(fa, fb), f"f differs after {done} steps"
                    da, db = a.diags(), b.diags()
                    for key in ("nrj", "nrjx", "nrjy", "mass", "l2norm"):
                        assert np.array_equal(da[key], db[key]), (key, done)
    finally:
        for m in ens_members + eager:
            m.close()


def test_ensemble_rejects_bad_requests(vlp):
    members = _make(vlp, 4)
    try:
        with pytest.raises(vlp.Vlp4dError):
            vlp.Ensemble([members[0], members[0]])
        with vlp.Ensemble(members) as ens:
            with pytest.raises(vlp.Vlp4dError):
                ens.step(5)  # beyond nbiter = 4
            ens.step(4)
        for m in members:
            assert m.iter == 4
    finally:
        for m in members:
            m.close()


@pytest.mark.parametrize("prec", ["FP64", "FP32_ACC"])
def test_autotune_keeps_results(vlp, tmp_path, prec):
    n = (64, 32, 32, 16)
    dt = 0.5 * (4 * math.pi / n[0]) / 6.0
    cfg = sld_box(n, vmax=6.0, dt=dt, nbiter=6)
    with vlp.Vlp4d(vlp.VlpConfig.make(*cfg), precision=getattr(vlp, prec)) as a, \
            vlp.Vlp4d(vlp.VlpConfig.make(*cfg), precision=getattr(vlp, prec)) as b:
        a.step(1)
        b.step(1)
        csv = tmp_path / "tune.csv"
        tvx, nb = a.autotune(str(csv))
        assert tvx in (2, 4, 8) and nb in (2, 3)
        assert (a.kernel_variants() >> 8) & 15 == tvx
        rows = csv.read_text().strip().splitlines()
        assert rows[0].startswith("kernel,") and len(rows) >= 2
        for fam in ("advect4d", "spline_vx", "spline_vy"):   # one default and one choice per tuned kernel
            fr = [r.split(",") for r in rows[1:] if r.split(",")[0] == fam]
            assert sum(int(r[-2]) for r in fr) == 1 and sum(int(r[-1]) for r in fr) == 1, fam
            dflt = next(float(r[-3]) for r in fr if r[-2] == "1")
            pick = next(float(r[-3]) for r in fr if r[-1] == "1")
            # another choice replaces the default only when >2 % faster, and
            # then it is the fastest
            assert pick == dflt or (pick < 0.98 * dflt and pick == min(float(r[-3]) for r in fr)), (fam, fr)
        a.step(5)
        b.step(5)
        # the advection tile leaves every cell's arithmetic unchanged; the
        # spline kernels split the line recursion differently (round-off,
        # then float rounding of the stored coefficients)
        bar = 1e-13 if prec == "FP64" else 1e-5
        assert relmax(a.get_f(), b.get_f()) < bar
        da, db = a.diags(), b.diags()
        assert relmax(da["nrj"], db["nrj"]) < bar


def test_ensemble_workload_48_members_match_oracle(vlp, orc):
    """The batched Landau workload itself: 48 members (16 x 32^4, 16 x 48^2 x
    64^2, 16 x 64^4), alpha ~ U[0.01, 0.5], k in {0.3, 0.5} (seeded), f stored
    in float with float interpolation arithmetic, autotuned 4-D advection
    tiles and spline kernels, 4 steps as two graph replays; every member
    against the FP64 oracle started from the member's own (float-rounded) f:
    f within 5e-5 (relative max-norm) and the field energy W = exp(2 nrj)
    within 1e-5 over the four steps -- the float-storage budget of
    test_gpu_fp32.py for a few steps."""
    steps = 4
    params = vlp.landau_ensemble_params(20261202)
    members, starts = [], []
    for n, alpha, k in params:
        cargs, f0 = vlp.landau_member_setup(n, alpha, k)
        sim = vlp.Vlp4d(vlp.VlpConfig.make(*cargs, steps), precision=vlp.FP32_ACC)
        sim.set_f(f0)
        sim.autotune()
        members.append(sim)
        starts.append(sim.get_f())
    with vlp.Ensemble(members) as ens:
        ens.step(steps)
    worst = {}
    for sim, f0, (n, alpha, k) in zip(members, starts, params):
        cargs, _ = vlp.landau_member_setup(n, alpha, k)
        st = orc.State(orc.make_config(*cargs, steps))
        st.set_f(f0)
        st.step(steps)
        ef = relmax(sim.get_f(), st.f_interior())
        w, wr = np.exp(2 * sim.diags()["nrj"][1:]), np.exp(2 * st.diags()["nrj"][1:])
        ew = relmax(w, wr)
        st.close()
        key = tuple(n)
        worst[key] = max(worst.get(key, (0, 0))[0], ef), max(worst.get(key, (0, 0))[1], ew)
        assert sim.iter == steps
        assert ef < 5e-5 and ew < 1e-5, (n, alpha, k, ef, ew)
    for key, (ef, ew) in worst.items():
        print(f"ensemble members {key}: worst f {ef:.2e}, worst W {ew:.2e}", flush=True)
    for sim in members:
        sim.close()
