"""Writes tests/golden/two_stream_64.npz: the oracle's run of the two-stream
workload (extension idcase 31) at 64^4, x, y in [0, 10 pi), v in [-8, 8],
dt = 0.05, 400 steps.

Contents: the five diagnostic series (nrj, nrjx, nrjy, mass, l2norm; 401
values each) and f subsampled every 8th point of every axis (8^4 values) at
steps 100, 200, 300 and 400.  The oracle is pinned to the reference phase by
phase (tests/test_oracle_*.py); a full run at this size takes minutes on the
CPU, too long for a GPU test to recompute, so the GPU test compares against
this file.

  OMP_NUM_THREADS=8 python tests/golden/make_two_stream_64.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

N, STEPS, DT = 64, 400, 0.05
SNAP = (100, 200, 300, 400)


def main():
    lx = 10 * math.pi
    c = oracle.make_config([N] * 4, [0.0, 0.0, -8.0, -8.0], [lx, lx, 8.0, 8.0], 31, DT, STEPS)
    st = oracle.State(c)
    snaps = []
    done = 0
    for s in SNAP:
        st.step(s - done)
        done = s
        snaps.append(st.f_interior()[::8, ::8, ::8, ::8].copy())
        print(f"step {s}", flush=True)
    d = st.diags()
    np.savez_compressed(os.path.join(HERE, "two_stream_64.npz"), steps=np.array(SNAP), f_sub=np.array(snaps),
                        **{k: d[k] for k in ("nrj", "nrjx", "nrjy", "mass", "l2norm")})
    st.close()


if __name__ == "__main__":
    main()
