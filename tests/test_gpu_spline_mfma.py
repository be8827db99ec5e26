"""The MFMA form of the vx and vy spline passes (VLP4D_SPLINE_MFMA=1 at
context creation): each pass as matrix products with the exact line
operator M = R E of the reference (VMO/Spline.hpp:114-199 on the alpha-sum
extension of VMO/Communication.hpp:756-898), on v_mfma_f64_16x16x4f64 for
double storage and v_mfma_f32_16x16x4f32 for float storage.

* FP64: the 4-D coefficients after computeCoeff_vxvy match the oracle at
  1e-13 and whole runs at the 1e-12 bar, on small non-cubic grids and with a
  128-point vx or vy axis (the register-resident operator of the 128-point
  kernel);
* float storage: the MFMA pass (float products and sums) against the scan
  pass (double recursion, float storage): the coefficient difference is
  float rounding (relative 1e-6), and a 20-step run stays within 1e-5.
"""
import math

import numpy as np
import pytest

from conftest import relmax, sld_box

pytestmark = pytest.mark.gpu

BIT_MFMA = 1 << 15


def _args(n, idcase=10, steps=4):
    lx, vmax = 4 * math.pi, 6.0
    dt = 0.5 * min(lx / n[0], lx / n[1]) / vmax
    return sld_box(n, vmax=vmax, dt=dt, nbiter=steps, idcase=idcase, lx=lx)


@pytest.mark.parametrize("n", [(16, 12, 20, 14), (24, 20, 16, 18), (16, 12, 128, 10), (16, 10, 12, 128)])
def test_mfma_spline_fp64_matches_oracle(vlp, orc, monkeypatch, n):
    monkeypatch.setenv("VLP4D_SPLINE_MFMA", "1")
    args = _args(n)
    gc, oc = vlp.VlpConfig.make(*args), orc.make_config(*args)
    rng = np.random.default_rng(5)
    h = orc.init_case(oc)
    orc.interior_of(h)[...] *= 1.0 + 0.1 * rng.standard_normal(orc.interior_shape(oc))
    f0 = np.ascontiguousarray(orc.interior_of(h))
    with vlp.Vlp4d(gc) as sim:
        assert sim.kernel_variants() & BIT_MFMA
        sim.set_f(f0)
        orc.exchange_halo(oc, h)
        orc.spline_xy(oc, h)
        fnp1 = h.copy()
        sim.spline_xy()
        sim.advect2d_poisson()
        orc.spline_vxvy(oc, fnp1)
        sim.spline_vxvy()
        err = relmax(sim.buffer(vlp.BUF_C4), orc.ext_view(fnp1))
        print(f"grid {n}: C4 vs oracle {err:.2e}", flush=True)
        assert err < 1e-13
    st = orc.State(oc)
    st.step(4)
    with vlp.Vlp4d(gc) as sim:
        sim.step(4)
        errf = relmax(sim.get_f(), st.f_interior())
        d = sim.diags()
    ref = st.diags()
    print(f"grid {n}: 4 steps f {errf:.2e}", flush=True)
    assert errf < 1e-12
    for k in ("nrj", "nrjx", "nrjy", "l2norm"):
        assert relmax(d[k], ref[k]) < 1e-12, k


@pytest.mark.parametrize("n", [(24, 20, 16, 18), (16, 12, 128, 10)])
def test_mfma_spline_float_storage_against_scan(vlp, monkeypatch, n):
    args = _args(n, steps=20)
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("VLP4D_SPLINE_MFMA", flag)
        with vlp.Vlp4d(vlp.VlpConfig.make(*args), precision=vlp.FP32_ACC) as sim:
            assert bool(sim.kernel_variants() & BIT_MFMA) == (flag == "1")
            sim.spline_xy()
            sim.advect2d_poisson()
            sim.spline_vxvy()
            c4 = sim.buffer(vlp.BUF_C4)
        with vlp.Vlp4d(vlp.VlpConfig.make(*args), precision=vlp.FP32_ACC) as sim:
            sim.step(20)
            out[flag] = (c4, sim.get_f(), sim.diags()["nrj"])
    e_c4 = relmax(out["1"][0], out["0"][0])
    e_f = relmax(out["1"][1], out["0"][1])
    e_w = relmax(np.exp(2 * out["1"][2]), np.exp(2 * out["0"][2]))
    print(f"grid {n} float storage: C4 {e_c4:.2e}, f after 20 steps {e_f:.2e}, W {e_w:.2e}", flush=True)
    assert e_c4 < 1e-6
    assert e_f < 1e-5 and e_w < 1e-5


@pytest.mark.parametrize("n", [(16, 12, 128, 10), (16, 10, 12, 128), (32, 24, 128, 128)])
def test_256_position_tiles_equal_128_position_tiles(vlp, monkeypatch, n):
    """Float storage: 256-position LDS tiles (VLP4D_SPLINE_WIDE=2, 1024 lanes)
    run the same line recursions as the 128-position tiles (four lanes of 32
    points per 128-point line), only grouped differently: bit-identical."""
    args = _args(n, steps=3)
    out = {}
    for wide in ("1", "2"):
        monkeypatch.setenv("VLP4D_SPLINE_WIDE", wide)
        with vlp.Vlp4d(vlp.VlpConfig.make(*args), precision=vlp.FP32_ACC) as sim:
            sim.step(3)
            out[wide] = (sim.get_f(), sim.diags()["nrj"])
    np.testing.assert_array_equal(out["1"][0], out["2"][0])
    np.testing.assert_array_equal(out["1"][1], out["2"][1])
