"""vlp_field_solve: the density, Poisson solve, field diagnostics and L2 norm
of a device-resident f through the C-ABI, against the oracle's field_rho,
field_poisson, diags_field and l2norm on the same f (VMO/Field.cpp:5-137,
VMO/Efield.cpp:3-106, VMO/Diags.cpp:16-76).

Bars: phi, Ex, Ey and nrj/nrjx/nrjy/l2norm at 1e-12 (relative max-norm);
mass = sum(phi) dx dy is round-off on both sides (phi has no k = 0 mode) and
is held to 16 eps sum|phi| dx dy (idcase 2 keeps rho, and its mass is held
at 1e-12).  f is the oracle's initial condition times
a seeded perturbation, on non-cubic grids; one grid has Nvy > 64 (the
density partials are then summed in slices first); idcase 2 takes the
reference's analytic field instead of the Poisson solve.  f_bits = 32 passes
f rounded to float, compared with the oracle on the same rounded values.
"""
import math

import numpy as np
import pytest

from conftest import relmax

pytestmark = pytest.mark.gpu

CASES = [((24, 20, 16, 18), 10, 4 * math.pi, 6.0), ((16, 12, 10, 72), 10, 4 * math.pi, 6.0),
         ((20, 24, 12, 14), 20, 10 * math.pi, 8.0), ((16, 16, 12, 10), 2, 1.0, 1.0)]


def oracle_field(orc, oc, f):
    h = orc.halo_zeros(oc)
    orc.interior_of(h)[...] = f
    phi, ex, ey = orc.field_poisson(oc, orc.field_rho(oc, h))
    d = orc.diags_field(oc, phi, ex, ey)
    return phi, ex, ey, np.array(list(d) + [orc.l2norm(oc, h)])


def run_field_solve(vlp, gc, f, bits, entry=None):
    from hip_buffers import DeviceArray
    vlp.load()  # the library (and its HIP runtime) first
    fd = DeviceArray.of(np.ascontiguousarray(f, dtype=np.float64 if bits == 64 else np.float32))
    n2 = gc.nxmax[0] * gc.nxmax[1]
    phi, ex, ey, diag = DeviceArray(8 * n2), DeviceArray(8 * n2), DeviceArray(8 * n2), DeviceArray(8 * 5)
    (entry or vlp.field_solve)(gc, fd.addr, bits, phi.addr, ex.addr, ey.addr, diag.addr)
    shape = (gc.nxmax[1], gc.nxmax[0])
    out = (phi.get(np.float64, shape), ex.get(np.float64, shape), ey.get(np.float64, shape), diag.get(np.float64, 5))
    for d in (fd, phi, ex, ey, diag):
        d.free()
    return out


@pytest.mark.parametrize("n,idcase,lx,vmax", CASES)
@pytest.mark.parametrize("bits", [64, 32])
def test_field_solve_matches_oracle(vlp, orc, n, idcase, lx, vmax, bits):
    lo = -lx / 2 if idcase == 2 else 0.0   # CYL02's bump sits around 0
    args = (list(n), [lo, lo, -vmax, -vmax], [lo + lx, lo + lx, vmax, vmax], idcase, 0.01, 1)
    gc, oc = vlp.VlpConfig.make(*args), orc.make_config(*args)
    rng = np.random.default_rng(sum(n) + idcase)
    f = orc.interior_of(orc.init_case(oc)) * (1.0 + 0.1 * rng.standard_normal(orc.interior_shape(oc)))
    if bits == 32:
        f = f.astype(np.float32).astype(np.float64)
    phi, ex, ey, d = run_field_solve(vlp, gc, f, bits)
    rphi, rex, rey, rd = oracle_field(orc, oc, f)
    assert relmax(phi, rphi) < 1e-12
    assert relmax(ex, rex) < 1e-12
    assert relmax(ey, rey) < 1e-12 or np.max(np.abs(rey)) == 0.0
    for k in (0, 1, 2, 4):
        assert abs(d[k] - rd[k]) <= 1e-12 * abs(rd[k]), k
    if idcase == 2:   # the analytic field keeps rho, whose sum is the mass
        assert abs(d[3] - rd[3]) <= 1e-12 * abs(rd[3])
    else:
        noise = 16 * 2.2e-16 * np.sum(np.abs(rphi)) * oc.dx[0] * oc.dx[1]
        assert abs(d[3]) <= noise and abs(rd[3]) <= noise


def test_field_solve_rejects_bad_arguments(vlp):
    gc = vlp.VlpConfig.make([16, 12, 10, 14], [0, 0, -6, -6], [4, 4, 6, 6], 10, 0.01, 1)
    with pytest.raises(vlp.Vlp4dError) as ei:
        vlp.field_solve(gc, 0, 64, 0, 0, 0, 0)
    assert ei.value.code == vlp.E_ARG
    with pytest.raises(vlp.Vlp4dError) as ei:
        vlp.field_solve(gc, 8, 16, 8, 8, 8, 8)
    assert ei.value.code == vlp.E_ARG
    bad = vlp.VlpConfig.make([8, 12, 10, 14], [0, 0, -6, -6], [4, 4, 6, 6], 10, 0.01, 1)
    with pytest.raises(vlp.Vlp4dError) as ei:
        vlp.field_solve(bad, 8, 64, 8, 8, 8, 8)
    assert ei.value.code == vlp.E_SIZE
