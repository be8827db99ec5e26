"""The oracle's whole step (A14).

The step is pinned as a composition: it must equal, bit for bit, the phase
functions (each pinned by its own test file) called in the order of
onetimestep (VMO/Timestep.hpp:34-97).  The reference pins no trajectory of
its own (no golden outputs ship with it), so the composed step is further
checked against what the Vlasov-Poisson system fixes:
* linear Landau damping of the SLD10 mode k = (1/2, 1/2): the electric energy
  oscillates at 2 Re(w) and decays at Im(w), w the root of
  1 + (1 + z Z(z)) / k^2 = 0, z = w / (sqrt2 k), Z(z) = i sqrt(pi) w(z)
  (Faddeeva function, scipy.special.wofz);
* the mass diagnostic is round-off (phi has no k = 0 mode) and the L2 norm of
  f is nearly conserved.
"""
import math

import numpy as np
import pytest
from scipy.optimize import fsolve
from scipy.special import wofz

from conftest import sld_box


def landau_root(k):
    def disp(p):
        z = (p[0] + 1j * p[1]) / (math.sqrt(2) * k)
        d = 1 + (1 + z * 1j * math.sqrt(math.pi) * wofz(z)) / k ** 2
        return [d.real, d.imag]
    return fsolve(disp, [1.5, -0.3])


def test_landau_damping_rate_and_frequency(orc):
    n, dt, steps = 24, 0.08, 125
    c = orc.make_config(*sld_box(n, vmax=6.0, dt=dt, nbiter=steps))
    st = orc.State(c)
    st.step(steps)
    d = st.diags()
    nrj = d["nrj"]
    t = np.arange(steps + 1) * dt
    peaks = [i for i in range(1, steps) if nrj[i] > nrj[i - 1] and nrj[i] >= nrj[i + 1]]
    assert len(peaks) >= 4
    rate = np.polyfit(t[peaks], nrj[peaks], 1)[0]
    wr, wi = landau_root(math.sqrt(0.5))
    assert abs(rate - wi) < 0.02 * abs(wi)
    # |E| peaks twice per oscillation period
    assert abs(np.mean(np.diff(t[peaks])) - math.pi / wr) < 0.05
    assert np.max(np.abs(d["mass"])) < 1e-15
    assert abs(d["l2norm"][-1] / d["l2norm"][0] - 1) < 1e-5


@pytest.mark.parametrize("idcase,n,lx,vmax,dt", [(10, (16, 12, 20, 14), 4 * math.pi, 6.0, 0.05),
                                                  (20, (14, 16, 12, 18), 10 * math.pi, 8.0, 0.1)])
def test_step_is_the_phases_in_timestep_order(orc, idcase, n, lx, vmax, dt):
    """vo_step equals the separately pinned phase functions composed in the
    order of onetimestep (VMO/Timestep.hpp:34-97), bit for bit: exchangeHalo
    and computeCoeff_xy on fn (:41-45), deep_copy fn -> fnp1 (:47),
    advect_2D_xy(dt/2) on fn (:49-51), field_rho / reduce / poisson (:53-63),
    computeCoeff_vxvy on fnp1 (:69-71), advect_4D(dt) with that field
    (:73-75), the field of fnp1 again with its diagnostics and L2 norm at
    index iter (:77-91; the half-step diagnostics at the same index are
    overwritten), then the swap (VMO/Vlp4d.cpp:45).  Two steps, so the second
    starts from the swapped buffer."""
    c = orc.make_config(*sld_box(n, vmax=vmax, dt=dt, nbiter=2, idcase=idcase, lx=lx))
    st = orc.State(c)
    h = st.fn()
    for it in (1, 2):
        orc.exchange_halo(c, h)
        orc.spline_xy(c, h)
        hp1 = h.copy()
        assert orc.advect_2d(c, h, 0.5 * dt) == 0
        phi, ex, ey = orc.field_poisson(c, orc.field_rho(c, h))
        orc.spline_vxvy(c, hp1)
        assert orc.advect_4d(c, ex, ey, hp1, dt) == 0
        phi2, ex2, ey2 = orc.field_poisson(c, orc.field_rho(c, hp1))
        dref = orc.diags_field(c, phi2, ex2, ey2)
        l2 = orc.l2norm(c, hp1)
        h = hp1
        st.step(1)
        assert st.iter == it
        np.testing.assert_array_equal(st.fn(), h)
        np.testing.assert_array_equal(st.field(0), phi2)
        np.testing.assert_array_equal(st.field(1), ex2)
        np.testing.assert_array_equal(st.field(2), ey2)
        d = st.diags()
        assert [d[k][it] for k in ("nrj", "nrjx", "nrjy", "mass")] == list(dref)
        # the L2 sum is an OpenMP reduction whose combine order may vary
        assert abs(d["l2norm"][it] - l2) <= 1e-15 * l2
    st.close()
