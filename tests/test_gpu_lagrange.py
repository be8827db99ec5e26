"""The Lagrange split-step solver (miniapps/vlp4d) through the C-ABI, against
the oracle (oracle/vlp4d_lag.c) at the 1e-12 bar of the FP64 path: initial
conditions of the four cases, multi-step runs on uneven grids (f and the
nrj/mass series, VO/diags.cpp:13-39), a step from a state set through
vlp_set_f, the 3-column nrj.out, and at 64^4 properties that hold at any
length (the total of f, conserved by every periodic 1-D advection with a
partition of unity; finite diagnostics).  Spline-only calls are refused.
"""
import math
import os

import numpy as np
import pytest

from conftest import relmax, sld_box

pytestmark = pytest.mark.gpu

CASES = [
    # (grid, idcase, vmax, box length, dt, steps)
    ((16, 12, 20, 14), 10, 6.0, 4 * math.pi, 0.1, 6),
    ((20, 16, 12, 18), 20, 8.0, 10 * math.pi, 0.1, 6),
    ((12, 14, 16, 10), 2, 1.0, 2.0, 0.05, 5),
    ((14, 12, 10, 16), 5, 1.0, 2.0, 0.05, 5),
    # rows of 128 and 80 points: two 64-lane chunks of x weights in the fused
    # pass, y weights wave-uniform (128) or per lane (80)
    ((128, 12, 10, 12), 10, 6.0, 4 * math.pi, 0.1, 3),
    ((80, 14, 12, 10), 20, 8.0, 10 * math.pi, 0.1, 3),
]


def _pair(vlp, orc, n, idcase, vmax, lx, dt, steps):
    args = sld_box(n, vmax=vmax, dt=dt, nbiter=steps, idcase=idcase, lx=lx)
    if idcase in (2, 5):  # the cylinder cases live on [-1, 1]^4
        args = (list(n), [-lx / 2, -lx / 2, -vmax, -vmax], [lx / 2, lx / 2, vmax, vmax], idcase, dt, steps)
    sim = vlp.Vlp4d(vlp.VlpConfig.make(*args), scheme=vlp.SCHEME_LAGRANGE)
    st = orc.LagState(orc.make_config(*args))
    return sim, st


@pytest.mark.parametrize("n,idcase,vmax,lx,dt,steps", CASES)
def test_lagrange_matches_oracle(vlp, orc, n, idcase, vmax, lx, dt, steps):
    sim, st = _pair(vlp, orc, n, idcase, vmax, lx, dt, steps)
    try:
        assert relmax(sim.get_f(), st.f()) < 1e-14
        sim.step(steps)
        st.step(steps)
        e = relmax(sim.get_f(), st.f())
        d, ref = sim.diags(), st.diags()
        print(f"idcase {idcase} {n} x{steps}: f {e:.2e} nrj {relmax(d['nrj'], ref['nrj']):.2e}")
        assert e < 1e-12
        if idcase == 5:
            # CYL05 is odd in vy: the initial density, so E at iteration 0,
            # is round-off on both sides (nrj = log ||E|| of noise, or the
            # reference's -1e9 floor); the y advection then builds a field
            assert d["nrj"][0] < math.log(1e-8) and ref["nrj"][0] < math.log(1e-8)
            assert relmax(d["nrj"][1:], ref["nrj"][1:]) < 1e-12
        else:
            assert relmax(d["nrj"], ref["nrj"]) < 1e-12
        # mass is the sum of phi (round-off of the field's size) or, for the
        # analytic fields, of rho
        scale = 1e-12 * max(1.0, float(np.max(np.abs(ref["mass"]))))
        assert np.max(np.abs(d["mass"] - ref["mass"])) < scale
        assert np.all(d["l2norm"] == 0)
    finally:
        sim.close()
        st.close()


def test_lagrange_set_f_then_step(vlp, orc):
    n, idcase, vmax, lx, dt, _ = CASES[0]
    sim, st = _pair(vlp, orc, n, idcase, vmax, lx, dt, 4)
    try:
        f = st.f() * (1 + 0.01 * np.random.default_rng(5).standard_normal(st.f().shape))
        sim.set_f(f)
        st.set_f(f)
        sim.step(2)
        st.step(2)
        assert relmax(sim.get_f(), st.f()) < 1e-12
    finally:
        sim.close()
        st.close()


def test_lagrange_save_and_refusals(vlp, tmp_path):
    n = (16, 12, 20, 14)
    args = sld_box(n, vmax=6.0, dt=0.1, nbiter=3, idcase=10, lx=4 * math.pi)
    with vlp.Vlp4d(vlp.VlpConfig.make(*args), scheme=vlp.SCHEME_LAGRANGE) as sim:
        sim.step(3)
        sim.save(str(tmp_path))
        rows = (tmp_path / "nrj.out").read_text().strip().splitlines()
        assert len(rows) == 4 and all(len(r.split()) == 3 for r in rows)
        assert not os.path.exists(tmp_path / "l2norm.out")
        d = sim.diags()
        assert abs(float(rows[2].split()[1]) - d["nrj"][2]) <= 1e-12 * abs(d["nrj"][2])
        for call in (lambda: sim.adjoint_step(np.zeros(n[::-1])), lambda: sim.autotune(None),
                     lambda: vlp.Ensemble([sim])):
            with pytest.raises(vlp.Vlp4dError):
                call()
        with pytest.raises(vlp.Vlp4dError):
            vlp.Vlp4d(vlp.VlpConfig.make(*args), precision=vlp.FP32, scheme=vlp.SCHEME_LAGRANGE)


def test_lagrange_64_conserves_the_total(vlp):
    steps = 40
    args = sld_box(64, vmax=6.0, dt=0.1, nbiter=steps, idcase=10, lx=4 * math.pi)
    with vlp.Vlp4d(vlp.VlpConfig.make(*args), scheme=vlp.SCHEME_LAGRANGE) as sim:
        s0 = float(sim.get_f().sum())
        sim.step(steps)
        s1 = float(sim.get_f().sum())
        d = sim.diags()
    print(f"Lagrange 64^4 x{steps}: total {s0:.15e} -> {s1:.15e}")
    assert abs(s1 - s0) < 1e-12 * abs(s0)
    assert np.all(np.isfinite(d["nrj"])) and np.max(np.abs(d["mass"])) < 1e-10


def test_lagrange_fused_and_separate_xy_passes_agree(vlp, orc, monkeypatch):
    # the x/2 + y/2 pair through one LDS plane (default where the plane fits)
    # and as two kernels (VLP4D_LAG_FUSED=0): same cells, same arithmetic;
    # the density is summed in another order
    n, idcase, vmax, lx, dt, steps = CASES[1]
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("VLP4D_LAG_FUSED", fused)
        sim, st = _pair(vlp, orc, n, idcase, vmax, lx, dt, steps)
        try:
            sim.step(steps)
            st.step(steps)
            assert relmax(sim.get_f(), st.f()) < 1e-12
            out[fused] = sim.get_f()
        finally:
            sim.close()
            st.close()
    assert relmax(out["1"], out["0"]) < 1e-13


def test_lagrange_driver_matches_oracle(orc, tmp_path):
    # the executable with --lagrange and the deck as its only positional
    # argument (the non-MPI miniapp's surface), against the oracle's series
    import subprocess
    from conftest import DECKS, ROOT
    txt = open(os.path.join(DECKS, "SLD10.dat")).read().replace(": 128", ": 16")
    txt = txt.replace("Time step : 0.01", "Time step : 0.1").replace("iterations : 40", "iterations : 4")
    (tmp_path / "d.dat").write_text(txt)
    p = subprocess.run([os.path.join(ROOT, "p3-miniapps_amd", "vlp4d"), "--lagrange", "d.dat"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert [ln for ln in p.stdout.splitlines() if ln.startswith("iter ")] == [f"iter {i}" for i in range(4)]
    assert "local domain" not in p.stdout
    rows = np.array([[float(v) for v in ln.split()] for ln in (tmp_path / "nrj.out").read_text().splitlines()])
    assert rows.shape == (5, 3) and not (tmp_path / "l2norm.out").exists()
    st = orc.LagState(orc.parse_deck(str(tmp_path / "d.dat")))
    st.step(4)
    ref = st.diags()
    st.close()
    np.testing.assert_allclose(rows[:, 1], ref["nrj"], rtol=1e-11)
    np.testing.assert_allclose(rows[:, 0], np.arange(5) * 0.1, rtol=1e-12)
