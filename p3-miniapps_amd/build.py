"""Build the in-tree HIP library and driver for gfx950.

Produces, next to this file:
  libvlp4d.so   the C-ABI library (include/vlp4d.h)
  vlp4d         the drop-in command-line driver

hipcc cross-compiles for gfx950 without a GPU.  Packed FP32 instructions are
disabled (-packed-fp32-ops) for every device compilation.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VLP4D_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}",
]

LIB = os.path.join(HERE, "libvlp4d.so")
EXE = os.path.join(HERE, "vlp4d")


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    srcs = [os.path.join(CSRC, n) for n in ("spline.hip", "spline_mfma.hip", "xy_fused.hip", "advect.hip", "advect_stream.hip", "adjoint.hip", "lagrange.hip", "field.hip", "field_solve.hip", "field_solve_noalloc.hip", "vlp4d_api.cpp")]
    objs = []
    procs = []
    for src in srcs:
        obj = os.path.join(CSRC, os.path.splitext(os.path.basename(src))[0] + ".o")
        cmd = [HIPCC, *COMMON, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), cmd))
        objs.append(obj)
    for p, cmd in procs:
        if p.wait() != 0:
            raise RuntimeError("compile failed: " + " ".join(cmd))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    _run([HIPCC, *COMMON, os.path.join(CSRC, "vlp4d_main.cpp"), "-o", EXE, f"-L{HERE}", "-lvlp4d",
          "-Wl,-rpath,$ORIGIN"])
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
