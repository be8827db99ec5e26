"""Whole-step properties of the oracle (A14) -- physics sanity, not parity.

The reference pins no trajectory of its own (no golden outputs ship with it),
so the full step is checked against what the Vlasov-Poisson system fixes:
* linear Landau damping of the SLD10 mode k = (1/2, 1/2): the electric energy
  oscillates at 2 Re(w) and decays at Im(w), w the root of
  1 + (1 + z Z(z)) / k^2 = 0, z = w / (sqrt2 k), Z(z) = i sqrt(pi) w(z)
  (Faddeeva function, scipy.special.wofz);
* the mass diagnostic is round-off (phi has no k = 0 mode) and the L2 norm of
  f is nearly conserved.
"""
import math

import numpy as np
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
