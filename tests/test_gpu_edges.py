"""Edge grid sizes against the oracles: the smallest the reference accepts
(N = MMAX + 2 = 10, VMO/Init.cpp:84-89), odd and prime sizes (the spline
line splits and tiles fall back to other kernels), and the largest (251 =
256 - 5, the reference's spline line buffer) along x and along vx, for the
spline step (1e-12, f and nrj) and the Lagrange split step.
"""
import math

import pytest

from conftest import relmax, sld_box

pytestmark = pytest.mark.gpu

EDGES = [
    ((10, 11, 13, 10), 10, 6.0, 4 * math.pi),
    ((251, 10, 10, 12), 20, 8.0, 10 * math.pi),
    ((12, 10, 251, 10), 10, 6.0, 4 * math.pi),
    ((10, 12, 10, 251), 20, 8.0, 10 * math.pi),
]


@pytest.mark.parametrize("n,idcase,vmax,lx", EDGES)
def test_spline_step_at_edge_sizes(vlp, orc, n, idcase, vmax, lx):
    steps = 3
    dt = 0.5 * (lx / max(n[0], n[1])) / vmax
    args = sld_box(n, vmax=vmax, dt=dt, nbiter=steps, idcase=idcase, lx=lx)
    st = orc.State(orc.make_config(*args))
    st.step(steps)
    with vlp.Vlp4d(vlp.VlpConfig.make(*args)) as sim:
        sim.step(steps)
        e = relmax(sim.get_f(), st.f_interior())
        en = relmax(sim.diags()["nrj"], st.diags()["nrj"])
        variants = sim.kernel_variants()
    st.close()
    print(f"spline {n}: f {e:.2e} nrj {en:.2e} variants {variants:#x}")
    assert e < 1e-12 and en < 1e-12


@pytest.mark.parametrize("n,idcase,vmax,lx", EDGES)
def test_lagrange_step_at_edge_sizes(vlp, orc, n, idcase, vmax, lx):
    steps = 3
    args = sld_box(n, vmax=vmax, dt=0.1, nbiter=steps, idcase=idcase, lx=lx)
    st = orc.LagState(orc.make_config(*args))
    st.step(steps)
    with vlp.Vlp4d(vlp.VlpConfig.make(*args), scheme=vlp.SCHEME_LAGRANGE) as sim:
        sim.step(steps)
        e = relmax(sim.get_f(), st.f())
        en = relmax(sim.diags()["nrj"], st.diags()["nrj"])
        fused = sim.kernel_variants() & 1
    st.close()
    print(f"lagrange {n}: f {e:.2e} nrj {en:.2e} fused {fused}")
    assert e < 1e-12 and en < 1e-12
