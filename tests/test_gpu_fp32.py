"""FP32 storage of f and the spline coefficients against the FP64 oracle.

Two modes store f and every spline coefficient array in float:
* FP32: the spline recursions, interpolation weights and stencil sums run in
  double, so the only single-precision error is the rounding of each stored
  array (f, the x/y/vx/vy coefficient passes);
* FP32_ACC: the interpolation weights and the stencil sums are float too.
Feet, the density reduction, the Poisson solve and the diagnostics are double
in both.  The bounds are each mode's stated error budget after the given
number of steps (not parity bars).  FP32 keeps the field-energy history
||E||_2 = exp(nrj) within 1e-6 relative of FP64 at every step; float weights
(their sum is 1 only to float rounding, the same bias every step for a given
foot fraction) and float sums drift linearly, which the looser FP32_ACC
bounds record.
"""
import math

import numpy as np
import pytest

from conftest import relmax, sld_box

pytestmark = pytest.mark.gpu

CASES = [
    (10, (32, 32, 32, 32), 40, 6.0, 4 * math.pi),
    (20, (48, 16, 48, 16), 20, 9.0, 8 * math.pi),
]
# mode -> (f relmax, |E| pointwise relative, L2 relative)
BUDGET = {"FP32": (5e-6, 1e-6, 1e-7), "FP32_ACC": (3e-5, 5e-5, 5e-6)}


@pytest.mark.parametrize("mode", ["FP32", "FP32_ACC"])
@pytest.mark.parametrize("idcase,n,steps,vmax,lx", CASES)
def test_fp32_storage_error_budget(vlp, orc, mode, idcase, n, steps, vmax, lx):
    dt = 0.5 * (lx / n[0]) / vmax
    args = sld_box(n, vmax=vmax, dt=dt, nbiter=steps, idcase=idcase, lx=lx)
    st = orc.State(orc.make_config(*args))
    st.step(steps)
    ref_f, ref_d = st.f_interior(), st.diags()
    with vlp.Vlp4d(vlp.VlpConfig.make(*args), precision=getattr(vlp, mode)) as sim:
        assert bool(sim.kernel_variants() & 64) == (mode == "FP32_ACC")
        sim.step(steps)
        f = sim.get_f()
        d = sim.diags()
    ef = relmax(f, ref_f)
    # pointwise relative error of ||E||_2 over the whole history
    e_nrj = float(np.max(np.abs(np.exp(d["nrj"] - ref_d["nrj"]) - 1)))
    e_l2 = float(np.max(np.abs(d["l2norm"] / ref_d["l2norm"] - 1)))
    print(f"{mode} vs fp64 oracle {n} x{steps}: f {ef:.2e}  |E| {e_nrj:.2e}  l2 {e_l2:.2e}")
    bf, be, bl = BUDGET[mode]
    assert ef < bf
    assert e_nrj < be
    assert e_l2 < bl
    # mass stays round-off: phi has no k = 0 mode in any precision
    phi = st.field(0)
    noise = 2.2e-16 * np.sum(np.abs(phi)) * lx / n[0] * lx / n[1]
    assert np.max(np.abs(d["mass"])) < 16 * noise


@pytest.mark.parametrize("mode", ["FP32", "FP32_ACC"])
def test_fp32_storage_on_a_128_point_x_axis(vlp, orc, mode):
    # N0 = 128 runs the y-streaming 4-D advection instantiated with the slab
    # row pitch 132 built in (one LDS base per pair of stencil rows) and the
    # fused x/y kernel on 128-point rows; same budgets as above
    n, steps, vmax, lx = (128, 16, 12, 10), 4, 6.0, 4 * math.pi
    dt = 0.5 * (lx / n[0]) / vmax
    args = sld_box(n, vmax=vmax, dt=dt, nbiter=steps, idcase=10, lx=lx)
    st = orc.State(orc.make_config(*args))
    st.step(steps)
    ref_f, ref_d = st.f_interior(), st.diags()
    with vlp.Vlp4d(vlp.VlpConfig.make(*args), precision=getattr(vlp, mode)) as sim:
        assert sim.kernel_variants() & (1 << 7)  # y-streaming 4-D advection
        assert sim.kernel_variants() & (1 << 12)  # fused x/y kernel
        sim.step(steps)
        f = sim.get_f()
        d = sim.diags()
    ef = relmax(f, ref_f)
    e_nrj = float(np.max(np.abs(np.exp(d["nrj"] - ref_d["nrj"]) - 1)))
    e_l2 = float(np.max(np.abs(d["l2norm"] / ref_d["l2norm"] - 1)))
    print(f"{mode} vs fp64 oracle {n} x{steps}: f {ef:.2e}  |E| {e_nrj:.2e}  l2 {e_l2:.2e}")
    bf, be, bl = BUDGET[mode]
    assert ef < bf
    assert e_nrj < be
    assert e_l2 < bl
