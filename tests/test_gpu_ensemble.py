"""Ensembles (one captured HIP graph per two steps of every member) and the
4-D advection autotuner, through the C-ABI.

The ensemble runs each member's kernels, in each member's order, from a
replayed graph; every replay is checked against a twin set of contexts
stepped eagerly with vlp_step, with eager work on the members between the
replays (an odd step count, which runs its last step eagerly and leaves the
members' buffers swapped, so the next replay needs a fresh capture).  The
autotuner changes only the 4-D advection's workgroup shape, which does not
change any cell's arithmetic: a tuned context gives the untuned one's result.
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
        assert sum(int(r.split(",")[-1]) for r in rows[1:]) == 1
        a.step(5)
        b.step(5)
        assert relmax(a.get_f(), b.get_f()) < 1e-14
        da, db = a.diags(), b.diags()
        assert relmax(da["nrj"], db["nrj"]) < 1e-14
