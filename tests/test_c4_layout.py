"""The C4 storage layout of the 4-D coefficients (vlp4d_internal.hpp,
c4_pitch / c4_elems), checked on the host through a probe built from the
library's own header.

Rows hold exactly the N0+4 x positions -2..N0+1 (no x halo), so the vy
spline pass writes contiguous runs; y, vx and vy keep their outer layer
(positions -3 and N+2), which the library zeroes once and no pass writes.
The checks: every position of the (N+6)^4 stencil range lands inside the
allocation; the coefficient positions (x in -2..N0+1, every other axis in
-3..N+2) map to distinct elements; the zero layer shares no element with
the coefficients the vy pass writes (-2..N+1 on every axis); the plane
pitch is a multiple of 16 elements (whole 64-byte blocks for float and
double)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "p3-miniapps_amd", "csrc")
GRIDS = [(12, 10, 8, 6), (16, 12, 20, 14), (128, 8, 4, 4), (10, 18, 6, 8), (48, 48, 6, 4)]


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("c4") / "c4_layout_probe")
    subprocess.run([hipcc, "-O1", "-std=c++17", "-I", CSRC, os.path.join(ROOT, "tests", "c4_layout_probe.cpp"),
                    "-o", exe], check=True, capture_output=True, timeout=300)

    def run(n):
        out = subprocess.run([exe, *map(str, n)], check=True, capture_output=True, text=True, timeout=60).stdout
        return tuple(int(v) for v in out.split())
    return run


def index(n, s, p0, p1, p2, p3):
    """Storage element of position (p0, p1, p2, p3): storage index p + 3 per axis."""
    s1, s2, s3 = s
    return (p0 + 3) + s1 * (p1 + 3) + s2 * (p2 + 3) + s3 * (p3 + 3)


@pytest.mark.parametrize("n", GRIDS)
def test_c4_layout(probe, n):
    s1, s2, s3, elems = probe(n)
    N0, N1, N2, N3 = n
    assert s1 == N0 + 4 and s2 % 16 == 0 and s2 > s1 * (N1 + 6) and s3 == s2 * (N2 + 6)
    ax = [np.arange(-3, N + 3) for N in n]
    # the whole stencil range -3..N+2 stays inside the storage
    lo = index(n, (s1, s2, s3), -3, -3, -3, -3)
    hi = index(n, (s1, s2, s3), N0 + 2, N1 + 2, N2 + 2, N3 + 2)
    assert lo == 0 and hi < elems
    # coefficient positions: x -2..N0+1, the other axes -3..N+2, all distinct
    P = np.meshgrid(np.arange(-2, N0 + 2), ax[1], ax[2], ax[3], indexing="ij")
    idx = index(n, (s1, s2, s3), *P)
    assert np.unique(idx).size == idx.size
    written = (P[1] <= N1 + 1) & (P[1] >= -2) & (P[2] <= N2 + 1) & (P[2] >= -2) & (P[3] <= N3 + 1) & (P[3] >= -2)
    assert not np.intersect1d(idx[written], idx[~written]).size
    # a vy-pass output run: the (x, y) positions -2..N+1 of one (vx, vy) are
    # one contiguous range of (N0+4)(N1+4) elements
    run = index(n, (s1, s2, s3), *np.meshgrid(np.arange(-2, N0 + 2), np.arange(-2, N1 + 2), 0, 0, indexing="ij"))
    assert np.array_equal(np.sort(run.ravel()), np.arange(run.min(), run.min() + run.size))
