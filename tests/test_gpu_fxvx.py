"""The fxvx diagnostics (vlp_write_fxvx and the driver's schedule).

The CSV files hold the cross sections f(x, y=0, vx, vy=Nvy/2) (one row per
vx) and, for the spline solver, f(x=0, y, vx=Nvx/2, vy) (one row per vy),
in the "%20.13le, " layout of the reference's print_fxvx
(VMO/Advection.hpp:500-590, miniapps/vlp4d/openmp/advect.hpp:318-341).  The
spline miniapp's own routine writes zeros (its guard is false at y index 0);
the library writes the sections it means, checked here against vlp_get_f.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import DECKS, ROOT, sld_box

pytestmark = pytest.mark.gpu


def _csv(path):
    return np.array([[float(v) for v in ln.split(",")] for ln in open(path).read().strip().splitlines()])


@pytest.mark.parametrize("prec,scheme", [("FP64", "SCHEME_SPLINE"), ("FP32", "SCHEME_SPLINE"),
                                         ("FP64", "SCHEME_LAGRANGE")])
def test_write_fxvx_sections(vlp, tmp_path, prec, scheme):
    n = (16, 12, 20, 14)
    cfg = vlp.VlpConfig.make(*sld_box(n, vmax=6.0, dt=0.05, nbiter=4, idcase=10, lx=4 * math.pi))
    with vlp.Vlp4d(cfg, precision=getattr(vlp, prec), scheme=getattr(vlp, scheme)) as sim:
        sim.step(2)
        sim.write_fxvx(str(tmp_path))
        f = sim.get_f()  # (vy, vx, y, x)
    fx = _csv(tmp_path / "fxvx_it000002.csv")
    assert fx.shape == (n[2], n[0])
    np.testing.assert_allclose(fx, f[n[3] // 2, :, 0, :], rtol=1e-12, atol=0)
    fy_path = tmp_path / "fyvy_it000002.csv"
    if scheme == "SCHEME_SPLINE":
        fy = _csv(fy_path)
        assert fy.shape == (n[3], n[1])
        np.testing.assert_allclose(fy, f[:, n[2] // 2, :, 0], rtol=1e-12, atol=0)
    else:
        assert not fy_path.exists()


@pytest.mark.parametrize("lagrange", [False, True])
def test_driver_writes_fxvx_on_schedule(tmp_path, lagrange):
    txt = open(os.path.join(DECKS, "SLD10.dat")).read().replace(": 128", ": 12")
    txt = txt.replace("Time step : 0.01", "Time step : 0.05").replace("iterations : 40", "iterations : 4")
    txt = txt.replace("Frequency of diagnostics : 8", "Frequency of diagnostics : 2")
    txt = txt.replace("Diagnostics of fxvx : 0", "Diagnostics of fxvx : 1")
    assert "fxvx : 1" in txt and "diagnostics : 2" in txt
    (tmp_path / "d.dat").write_text(txt)
    args = [os.path.join(ROOT, "p3-miniapps_amd", "vlp4d")] + (["--lagrange", "d.dat"] if lagrange else ["-f", "d.dat"])
    p = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    its = [0, 2, 4] if lagrange else [2, 4]
    assert [ln for ln in p.stdout.splitlines() if ln.startswith("print_fxvx")] == [f"print_fxvx {i}" for i in its]
    d = tmp_path / "data" / ("vlp4d" if lagrange else "vlp4d_mpi")
    names = sorted(os.listdir(d))
    want = sorted([f"fxvx_it{i:06d}.csv" for i in its] + ([] if lagrange else [f"fyvy_it{i:06d}.csv" for i in its]))
    assert names == want
    assert np.any(_csv(d / f"fxvx_it{its[-1]:06d}.csv") != 0)
