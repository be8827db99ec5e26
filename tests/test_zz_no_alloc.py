"""vlp_field_solve_noalloc (the field solve with its scratch in the library's
static device array, no allocation) against vlp_field_solve (the same
kernels on scratch it allocates per call), through the C-ABI.

Two calls in a row, the second on a smaller grid (the array then holds the
first call's larger contents, which no kernel may read), each compared with
the allocating entry point on the same input: the same kernels on the same
data, so the results must agree bit for bit.  A grid whose scratch exceeds
the array is refused with VLP_E_SIZE before any launch.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(vlp, entry, gc, f, bits):
    from hip_buffers import DeviceArray, hip
    vlp.load()
    fd = DeviceArray.of(np.ascontiguousarray(f, dtype=np.float64 if bits == 64 else np.float32))
    n2 = gc.nxmax[0] * gc.nxmax[1]
    outs = [DeviceArray(8 * n2), DeviceArray(8 * n2), DeviceArray(8 * n2), DeviceArray(8 * 5)]
    entry(gc, fd.addr, bits, *[o.addr for o in outs])
    assert hip().hipDeviceSynchronize() == 0
    shape = (gc.nxmax[1], gc.nxmax[0])
    res = [outs[0].get(np.float64, shape), outs[1].get(np.float64, shape), outs[2].get(np.float64, shape),
           outs[3].get(np.float64, 5)]
    for d in [fd] + outs:
        d.free()
    return res


@pytest.mark.parametrize("bits", [64, 32])
def test_noalloc_matches_allocating_entry_point(vlp, orc, bits):
    rng = np.random.default_rng(99 + bits)
    cases = [((128, 96, 20, 72), 10, 4 * math.pi, 6.0),   # Nvy > 64: slice sums of the density partials
             ((24, 20, 16, 18), 20, 10 * math.pi, 8.0)]
    for n, idcase, lx, vmax in cases:
        args = (list(n), [0.0, 0.0, -vmax, -vmax], [lx, lx, vmax, vmax], idcase, 0.01, 1)
        gc, oc = vlp.VlpConfig.make(*args), orc.make_config(*args)
        f = orc.interior_of(orc.init_case(oc)) * (1.0 + 0.1 * rng.standard_normal(orc.interior_shape(oc)))
        ref = _run(vlp, vlp.field_solve, gc, f, bits)
        got = _run(vlp, vlp.field_solve_noalloc, gc, f, bits)
        print(f"grid {n} f_bits {bits}: nrj {got[3][0]:.15e} (allocating {ref[3][0]:.15e})", flush=True)
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)
        assert np.all(np.isfinite(got[3]))


def test_noalloc_refuses_grids_beyond_its_scratch(vlp):
    vlp.load()
    # scratch ~ Nx*Ny*Nvy doubles: 250 x 250 x 250 is far beyond 3 * 2^20
    gc = vlp.VlpConfig.make([250, 250, 12, 250], [0, 0, -6, -6], [4, 4, 6, 6], 10, 0.01, 1)
    with pytest.raises(vlp.Vlp4dError) as ei:
        vlp.field_solve_noalloc(gc, 8, 64, 8, 8, 8, 8)   # refused before any pointer is used
    assert ei.value.code == vlp.E_SIZE
