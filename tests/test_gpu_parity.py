"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle.

Each phase of onetimestep (VMO/Timestep.hpp:34-97) is fed the same input on
both sides and compared in relative max-norm (max|gpu-oracle| / max|oracle|).
The bar is 1e-12 (relative max-norm) for FP64; phase-level checks use 1e-13
(single applications; differences are FMA contraction and summation order).
The grids are deliberately non-cubic so that a swapped dimension or stride
cannot pass.
"""
import math

import numpy as np
import pytest

from conftest import relmax, sld_box

pytestmark = pytest.mark.gpu

# (Nx, Ny, Nvx, Nvy) non-uniform; CFL-valid dt (dt*V <= dx).
# The first two run every spline pass with one lane per line; the third puts
# 2-4 lanes on each line (x: 4 lanes of 10, y/vx/vy: 2 lanes) like 64^4 and
# 128^4 do.
SHAPES = [(16, 12, 20, 14), (24, 20, 16, 18), (40, 36, 48, 34)]


def _configs(vlp, orc, n, idcase=10, dt=0.05, nbiter=8, vmax=6.0, lx=4 * math.pi):
    args = sld_box(n, vmax=vmax, dt=dt, nbiter=nbiter, idcase=idcase, lx=lx)
    return vlp.VlpConfig.make(*args), orc.make_config(*args)


def _cxy_view(orc, h):
    # x,y on -2..N+1, vx,vy interior
    return h[3:-3, 3:-3, 1:-1, 1:-1]


@pytest.mark.parametrize("n", SHAPES)
def test_init_matches_oracle(vlp, orc, n):
    gc, oc = _configs(vlp, orc, n)
    with vlp.Vlp4d(gc) as sim:
        f = sim.get_f()
    ref = orc.interior_of(orc.init_case(oc))
    assert relmax(f, ref) < 1e-14


@pytest.mark.parametrize("n", SHAPES)
def test_phases_match_oracle(vlp, orc, n):
    gc, oc = _configs(vlp, orc, n)
    rng = np.random.default_rng(1234)
    h = orc.init_case(oc)
    # perturb so that no symmetry hides an index error
    orc.interior_of(h)[...] *= 1.0 + 0.1 * rng.standard_normal(orc.interior_shape(oc))
    f0 = np.ascontiguousarray(orc.interior_of(h))
    with vlp.Vlp4d(gc) as sim:
        sim.set_f(f0)
        # exchangeHalo + computeCoeff_xy
        orc.exchange_halo(oc, h)
        orc.spline_xy(oc, h)
        fnp1 = h.copy()
        sim.spline_xy()
        cxy = sim.buffer(vlp.BUF_CXY)
        assert relmax(cxy, _cxy_view(orc, h)) < 1e-13
        # advect_2D_xy(dt/2) + field_rho/reduce/poisson
        assert orc.advect_2d(oc, h, 0.5 * oc.dt) == 0
        rho = orc.field_rho(oc, h)
        phi, ex, ey = orc.field_poisson(oc, rho)
        sim.advect2d_poisson()
        assert relmax(sim.buffer(vlp.BUF_EX), ex) < 1e-12
        assert relmax(sim.buffer(vlp.BUF_EY), ey) < 1e-12
        assert relmax(sim.buffer(vlp.BUF_RHO), phi) < 1e-12
        # computeCoeff_vxvy
        orc.spline_vxvy(oc, fnp1)
        sim.spline_vxvy()
        assert relmax(sim.buffer(vlp.BUF_C4), orc.ext_view(fnp1)) < 1e-13
        # advect_4D + field + diags (oracle uses its own E)
        assert orc.advect_4d(oc, ex, ey, fnp1, oc.dt) == 0
        sim.advect4d_poisson()
        assert relmax(sim.get_f(), orc.interior_of(fnp1)) < 1e-12
        rho2 = orc.field_rho(oc, fnp1)
        phi2, ex2, ey2 = orc.field_poisson(oc, rho2)
        d = orc.diags_field(oc, phi2, ex2, ey2)
        g = sim.diags()
        assert abs(g["nrj"][1] - d[0]) <= 1e-12 * abs(d[0])
        assert abs(g["nrjx"][1] - d[1]) <= 1e-12 * abs(d[1])
        assert abs(g["nrjy"][1] - d[2]) <= 1e-12 * abs(d[2])
        noise = 2.2e-16 * np.sum(np.abs(phi2)) * oc.dx[0] * oc.dx[1]
        assert abs(g["mass"][1]) < 16 * noise and abs(d[3]) < 16 * noise
        l2 = orc.l2norm(oc, fnp1)
        assert abs(g["l2norm"][1] - l2) <= 1e-12 * l2


@pytest.mark.parametrize("idcase,n,steps,vmax,lx", [
    (10, (16, 16, 16, 16), 20, 6.0, 4 * math.pi),
    (20, (20, 12, 24, 16), 12, 9.0, 10 * math.pi),
    (20, (64, 64, 64, 64), 3, 8.0, 10 * math.pi),
])
def test_multistep_matches_oracle(vlp, orc, idcase, n, steps, vmax, lx):
    dt = 0.5 * (lx / n[0]) / vmax
    gc, oc = _configs(vlp, orc, n, idcase=idcase, dt=dt, nbiter=steps, vmax=vmax, lx=lx)
    st = orc.State(oc)
    st.step(steps)
    ref_f = st.f_interior()
    ref_d = st.diags()
    with vlp.Vlp4d(gc) as sim:
        sim.step(steps)
        f = sim.get_f()
        d = sim.diags()
    assert relmax(f, ref_f) < 1e-12
    for k in ("nrj", "nrjx", "nrjy", "l2norm"):
        assert relmax(d[k], ref_d[k]) < 1e-12, k
    # mass = sum(phi) dx dy is round-off by construction (phi has no k=0
    # mode, VMO/Efield.cpp:38): both sides are noise of size eps*sum|phi|dxdy.
    phi = st.field(0)
    noise = 2.2e-16 * np.sum(np.abs(phi)) * oc.dx[0] * oc.dx[1]
    assert np.max(np.abs(d["mass"])) < 16 * noise
    assert np.max(np.abs(ref_d["mass"])) < 16 * noise


def test_cfl_violation_reports_error(vlp):
    # dt*V/dx = 3 > 1: the reference prints "Time step is too large" and exits.
    args = sld_box(16, vmax=6.0, dt=3 * (4 * math.pi / 16) / 6.0, nbiter=2)
    with vlp.Vlp4d(vlp.VlpConfig.make(*args)) as sim:
        with pytest.raises(vlp.Vlp4dError) as ei:
            sim.step(1)
        assert ei.value.code == vlp.E_CFL


@pytest.mark.parametrize("precision", ["FP64", "FP32_ACC"])
def test_multistep_call_equals_single_steps(vlp, precision):
    # vlp_step(n > 1) queues step s + 1 before it reads step s's error flag
    # and phase events; f, the diagnostics and the timer counts must be
    # those of n calls of one step.
    args = sld_box((24, 20, 16, 18), vmax=6.0, dt=0.05, nbiter=7)
    prec = getattr(vlp, precision)
    with vlp.Vlp4d(vlp.VlpConfig.make(*args), precision=prec) as a, \
            vlp.Vlp4d(vlp.VlpConfig.make(*args), precision=prec) as b:
        a.step(7)
        for _ in range(7):
            b.step(1)
        assert a.iter == b.iter == 7
        assert np.array_equal(a.get_f(), b.get_f())
        da, db = a.diags(), b.diags()
        for k in da:
            assert np.array_equal(da[k], db[k]), k
        ta, tb = a.timers(), b.timers()
        assert all(ta[k][1] == tb[k][1] for k in ta)
        assert ta["MainLoop"][1] == 7 and ta["advec4D"][0] > 0


def test_cfl_violation_in_a_multistep_call(vlp):
    # the failing step is the first of three: the error names it, iter stays
    # at that step and the context refuses further steps
    args = sld_box(16, vmax=6.0, dt=3 * (4 * math.pi / 16) / 6.0, nbiter=3)
    with vlp.Vlp4d(vlp.VlpConfig.make(*args)) as sim:
        with pytest.raises(vlp.Vlp4dError) as ei:
            sim.step(3)
        assert ei.value.code == vlp.E_CFL
        assert sim.iter == 1
        with pytest.raises(vlp.Vlp4dError):
            sim.step(1)


# The vx-blocked 4-D advection (taken when the y-streaming kernel does not
# fit the geometry, forced here with VLP4D_YSTREAM=0) has tuned variants picked
# at launch from the environment (cells per lane block, E rows staged in LDS
# when Nx % 64 == 0, XCD-aware workgroup order).  Every combination must give
# the oracle's result; the first shape takes the LDS path, the second cannot.
@pytest.mark.parametrize("n", [(64, 20, 24, 16), (40, 36, 48, 34)])
@pytest.mark.parametrize("L,elds,remap", [(4, 1, 1), (4, 0, 0), (2, 1, 0), (2, 0, 1)])
def test_advect4d_variants_match_oracle(vlp, orc, monkeypatch, n, L, elds, remap):
    monkeypatch.setenv("VLP4D_YSTREAM", "0")
    monkeypatch.setenv("VLP4D_ADVECT_L", str(L))
    monkeypatch.setenv("VLP4D_ELDS", str(elds))
    monkeypatch.setenv("VLP4D_XCD_REMAP", str(remap))
    steps = 2
    dt = 0.5 * (4 * math.pi / n[0]) / 6.0
    gc, oc = _configs(vlp, orc, n, dt=dt, nbiter=steps)
    st = orc.State(oc)
    st.step(steps)
    with vlp.Vlp4d(gc) as sim:
        bits = sim.kernel_variants()
        sim.step(steps)
        f = sim.get_f()
        d = sim.diags()
    assert bool(bits & 8) == (L == 4)
    assert bool(bits & 32) == (remap == 1)
    assert bool(bits & 16) == (elds == 1 and n[0] % 64 == 0)
    assert relmax(f, st.f_interior()) < 1e-12
    assert relmax(d["nrj"], st.diags()["nrj"]) < 1e-12


# The y-streaming 4-D advection: a lane adds the plane to its four open cells
# in one pass over the LDS rows when their stencil offsets agree, in one pass
# per distinct offset pattern otherwise; feet outside the streamed pattern
# (here: the vy = 0 slab, whose y offset follows the sign of E) take the
# direct gather.  VLP4D_YS_FORCE makes every lane take one pass per cell (1),
# every foot the direct gather (2), or both (3); each path must give the
# oracle's result, with two LDS plane buffers (the next plane's loads overlap
# the step) and with three (loads two steps ahead), at the default vx tile
# and at tile 4.  Shapes: a vy = 0 slab (Nvy even), Nx a multiple of 16 but
# not of 64, several waves per workgroup and one; the last has 8 vx tiles, so
# its workgroups take the XCD-grouped (tile, vy) order.
@pytest.mark.parametrize("n", [(32, 12, 20, 14), (48, 20, 16, 18), (64, 16, 32, 12), (32, 12, 32, 16)])
@pytest.mark.parametrize("force", [0, 1, 2, 3])
@pytest.mark.parametrize("nbuf", [2, 3])
@pytest.mark.parametrize("tvx", [0, 4])
def test_advect4d_ystream_paths_match_oracle(vlp, orc, monkeypatch, n, force, nbuf, tvx):
    monkeypatch.setenv("VLP4D_YS_FORCE", str(force))
    monkeypatch.setenv("VLP4D_YS_NBUF", str(nbuf))
    if tvx:
        monkeypatch.setenv("VLP4D_YS_TVX", str(tvx))
    steps = 2
    dt = 0.5 * (4 * math.pi / n[0]) / 6.0
    gc, oc = _configs(vlp, orc, n, dt=dt, nbiter=steps)
    st = orc.State(oc)
    st.step(steps)
    with vlp.Vlp4d(gc) as sim:
        bits = sim.kernel_variants()
        sim.step(steps)
        f = sim.get_f()
        d = sim.diags()
    assert bits & 128, "geometry expected to take the y-streaming kernel"
    if tvx:
        assert (bits >> 8) & 15 == tvx
    assert relmax(f, st.f_interior()) < 1e-12
    for k in ("nrj", "nrjx", "nrjy", "l2norm"):
        assert relmax(d[k], st.diags()[k]) < 1e-12, k


@pytest.mark.parametrize("precision", ["FP32", "FP32_ACC"])
def test_advect4d_ystream_fp32_matches_blocked(vlp, monkeypatch, precision):
    """Float storage: the y-streaming kernel and the vx-blocked kernel compute
    the same interpolation in the same arithmetic type; their results agree to
    the rounding of that type (summation order differs)."""
    n = (32, 16, 24, 14)
    dt = 0.5 * (4 * math.pi / n[0]) / 6.0
    gc = vlp.VlpConfig.make(*sld_box(n, vmax=6.0, dt=dt, nbiter=3))
    out = {}
    for ys in ("1", "0"):
        monkeypatch.setenv("VLP4D_YSTREAM", ys)
        with vlp.Vlp4d(gc, precision=getattr(vlp, precision)) as sim:
            assert bool(sim.kernel_variants() & 128) == (ys == "1")
            sim.step(3)
            out[ys] = (sim.get_f(), sim.diags())
    # a float ulp is 6e-8 relative; float sums in a different order differ by
    # a few ulp per step
    tol = 4e-7 if precision == "FP32" else 2e-6
    assert relmax(out["1"][0], out["0"][0]) < tol
    assert relmax(out["1"][1]["nrj"], out["0"][1]["nrj"]) < 1e-6


# The fused x/y spline + 2-D advection kernel (one (vx, vy) plane per LDS
# pass, density of the half step summed over groups of vx planes) against the
# oracle's computeCoeff_xy, advect_2D_xy and field solve, then whole steps.
# Shapes take both chunk sizes (C = 12: Nx = 48, Ny = 24; C = 16) and uneven
# lane counts per line; VLP4D_XY_FUSED=0 must give the same result through the
# separate kernels.
XY_SHAPES = [(48, 24, 20, 14), (32, 16, 24, 18), (64, 32, 16, 12)]


@pytest.mark.parametrize("n", XY_SHAPES)
def test_xy_fused_phases_match_oracle(vlp, orc, n):
    gc, oc = _configs(vlp, orc, n)
    rng = np.random.default_rng(4321)
    h = orc.init_case(oc)
    orc.interior_of(h)[...] *= 1.0 + 0.1 * rng.standard_normal(orc.interior_shape(oc))
    f0 = np.ascontiguousarray(orc.interior_of(h))
    with vlp.Vlp4d(gc) as sim:
        assert sim.kernel_variants() & 4096, "geometry expected to take the fused xy kernel"
        sim.set_f(f0)
        orc.exchange_halo(oc, h)
        orc.spline_xy(oc, h)
        sim.spline_xy()
        assert relmax(sim.buffer(vlp.BUF_CXY), _cxy_view(orc, h)) < 1e-13
        assert orc.advect_2d(oc, h, 0.5 * oc.dt) == 0
        rho = orc.field_rho(oc, h)
        phi, ex, ey = orc.field_poisson(oc, rho)
        sim.advect2d_poisson()
        assert relmax(sim.buffer(vlp.BUF_EX), ex) < 1e-12
        assert relmax(sim.buffer(vlp.BUF_EY), ey) < 1e-12
        assert relmax(sim.buffer(vlp.BUF_RHO), phi) < 1e-12


@pytest.mark.parametrize("n", XY_SHAPES)
@pytest.mark.parametrize("fused", ["1", "0"])
def test_xy_fused_multistep_matches_oracle(vlp, orc, monkeypatch, n, fused):
    monkeypatch.setenv("VLP4D_XY_FUSED", fused)
    steps = 3
    dt = 0.5 * (4 * math.pi / n[0]) / 6.0
    gc, oc = _configs(vlp, orc, n, dt=dt, nbiter=steps)
    st = orc.State(oc)
    st.step(steps)
    with vlp.Vlp4d(gc) as sim:
        assert bool(sim.kernel_variants() & 4096) == (fused == "1")
        sim.step(steps)
        f = sim.get_f()
        d = sim.diags()
    assert relmax(f, st.f_interior()) < 1e-12
    for k in ("nrj", "nrjx", "nrjy", "l2norm"):
        assert relmax(d[k], st.diags()[k]) < 1e-12, k
