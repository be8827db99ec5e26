"""Benchmark of the vlp4d spline time step on one MI355X.

Workload: the reference's wk/SLD10.dat case at 128^4 (Landau damping,
idcase 10, x,y in [0, 12.56637112), v in [-9, 9], dt = 0.01), synthetic
initial data generated on the GPU.  A step is one onetimestep
(miniapps/vlp4d_mpi/openmp/Timestep.hpp:34-97) including both Poisson solves
and the diagnostics -- the reference's MainLoop timer scope.

Metric: 4-D cell updates per second = N^4 * steps / wall time of the timed
steps (inputs resident in HBM).  Prints ONE JSON line.  By default the same
run times the FP32-storage step (f and coefficients stored in float, float
interpolation arithmetic, FP64 field solve / density / diagnostics) and the
plain FP64 step: `value` is the FP32-storage rate, the FP64 rate and the
speed-up sit beside it, with each mode's roofline and the field-energy
history of the FP32 run against the FP64 run over the same steps.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision 64|32|33]
                  [--n 128] [--no-cpu-baseline]   # --precision: one mode only
  python bench.py --lagrange [--n 128] # the Lagrange split-step solver (miniapps/vlp4d)
  python bench.py --calibrate-copy     # plain copy kernel of known size, for
                                       # scaling rocprofv3 FETCH_SIZE/WRITE_SIZE
"""
import argparse
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))

METRIC = "4D cell updates/s at 128^4 (FP32-store vs FP64 GPU) + % of HBM roofline"
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TFS = 78.6      # FP64 vector peak of MI355X (also its non-packed FP32 rate)
VALU_PEAK_FP32_TFS = 157.3  # FP32 vector peak, reached only with packed FP32 (disabled in this build)
# Algorithmic flops of one interp_4D + computeFeet per cell, counted from the
# reference loops (VMO/Advection.hpp:166-218, 221-296): 336 FMA + 16 MUL in
# the contraction, 56 for the 4x4 weights, ~100 for the feet.
ADVECT4D_FLOP_PER_CELL = 860
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "advect4d_traffic_sld10_128.txt")
LAG_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "lagrange_traffic_sld10_128.txt")
A100_REF = 2.38e9  # 40*128^4 / 4.51995 s, A100 stdpar MainLoop (docs/vlp4d.md:148,168-171)
SLD10 = dict(minphy=[0.0, 0.0, -9.0, -9.0], maxphy=[12.56637112, 12.56637112, 9.0, 9.0], idcase=10, dt=0.01)

DECK_TEMPLATE = """** Definition of mesh
Number of points in Ex with the coarse mesh : {n}
Number of points in Ey with the coarse mesh : {n}
Number of points in Vx with the coarse mesh : {n}
Number of points in Vy with the coarse mesh : {n}

** Defintion of the geometry of the domain
Minimal value of Ex : 0.
Maximal value of Ex : 12.56637112
Minimal value of Ey : 0.
Maximal value of Ey : 12.56637112

Minimal value of Vx : -9.
Maximal value of Vx :  9.
Minimal value of Vy : -9.
Maximal value of Vy :  9.

** Considered test cases
 -10- Landau Damping
 -11- Landau Damping 2
 -20- Two beam instability
Number of the chosen test case : 10

** Iterations in time and diagnostics
Time step : 0.01
Number of total iterations : {nbiter}
Frequency of diagnostics : 8
Diagnostics of fxvx : 0
"""


def load_pkg():
    import importlib.util
    name = "p3_miniapps_amd"
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "p3-miniapps_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def algorithmic_bytes(n, b):
    """Minimal HBM traffic of one step (bytes): read f, write CXY; read CXY
    (advect2D); read CXY, write C4 (vx,vy splines); read C4, write f^{n+1}."""
    e = n + 4
    return b * (2 * n ** 4 + 3 * e * e * n * n + 2 * e ** 4)


def advect4d_bytes(n, b):
    """Algorithmic bytes of one advect_4D launch: read C4 once, write f^{n+1}."""
    return b * ((n + 4) ** 4 + n ** 4)


def measured_traffic(kernel, precision, n, path=TRAFFIC_FILE):
    """HBM bytes per launch of a kernel from a committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration),
    or None when no summary for this kernel, precision and size exists."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"^TRAFFIC kernel=(\w+) precision=(\d+) n=(\d+) bytes_per_launch=([0-9.eE+]+)", line)
        if m and m.group(1) == kernel and int(m.group(2)) == precision and int(m.group(3)) == n:
            return float(m.group(4))
    return None


def calibrate_copy():
    """A plain copy kernel of known size (torch's elementwise copy of 2^27
    doubles: 1 GiB read + 1 GiB written per launch), run 3 times."""
    import torch
    x = torch.ones(2 ** 27, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 3
    print(json.dumps({"copy_kernel": "torch.mul(x, 1.0, out=y)", "read_bytes": 2 ** 30, "write_bytes": 2 ** 30,
                      "launches": 3, "GBps": 2 * 2 ** 30 / dt / 1e9}))


def cpu_model():
    """CPU model and physical cores per socket from lscpu (None if absent)."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return None, None
    model = re.search(r"^Model name:\s*(.+)$", out, re.M)
    cps = re.search(r"^Core\(s\) per socket:\s*(\d+)", out, re.M)
    return (model.group(1).strip() if model else None), (int(cps.group(1)) if cps else None)


def cpu_baseline(n, warmup=1, steps=3):
    """Oracle (C/OpenMP port of the reference OpenMP step) on the host cores:
    `warmup` untimed + `steps` timed MainLoop steps of the same deck, per-step
    times printed by the oracle driver (-T).  Threads: the physical cores of
    one socket, at most 16 (this host's share of cores per GPU)."""
    exe = os.path.join(ROOT, "oracle", "vlp4d_oracle")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    model, per_socket = cpu_model()
    cores = min(16, per_socket or os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
    with tempfile.TemporaryDirectory() as d:
        deck = os.path.join(d, "bench.dat")
        with open(deck, "w") as fh:
            fh.write(DECK_TEMPLATE.format(n=n, nbiter=warmup + steps))
        t0 = time.time()
        out = subprocess.run([exe, "-f", deck, "-T"], cwd=d, env=env, capture_output=True, text=True,
                             check=True).stdout
        wall = time.time() - t0
    per = [float(t) for _, t in re.findall(r"^steptime (\d+) ([0-9.eE+-]+)", out, re.M)]
    timed = per[warmup:]
    sec = sum(timed) / len(timed)
    return {"value": n ** 4 / sec, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "cpu_model": model, "cores_per_socket": per_socket,
            "sample": f"{warmup} warm-up + {steps} timed MainLoop steps of the same SLD10 deck at {n}^4, "
                      f"oracle/vlp4d_oracle (FP64, OpenMP {cores} threads, {sec:.2f} s/step, "
                      f"steps {', '.join(f'{t:.2f}' for t in per)} s, {wall:.1f} s incl. init)"}


def ensemble_bench(pkg, steps, seed=20261202):
    """Batched ensemble of independent Landau solves (one captured HIP graph
    per two steps of every member): 16 at 32^4, 16 at 48^2 x 64^2, 16 at
    64^4, f stored in float with float interpolation arithmetic, perturbation
    amplitude alpha ~ U[0.01, 0.5] and wave number k in {0.3, 0.5} from a
    seeded generator, box [0, 2 pi / k)^2 x [-6, 6]^2, dt = dx / (2 V),
    autotuned 4-D advection tiles and spline kernels.  Reports solves per
    second (a solve = the given number of steps of one member) and the same
    members stepped one after another for comparison."""
    import numpy as np
    members, cfgs = [], []
    for n, alpha, k in pkg.landau_ensemble_params(seed):
        cargs, f0 = pkg.landau_member_setup(n, alpha, k)
        cfg = pkg.VlpConfig.make(*cargs, 2 * steps + 2)
        sim = pkg.Vlp4d(cfg, precision=pkg.FP32_ACC)
        sim.set_f(f0)
        sim.autotune()
        members.append(sim)
        cfgs.append((n, alpha, k))
    shapes = [c[0] for c in cfgs]
    # the same members stepped one after another (eager), then the ensemble
    t0 = time.perf_counter()
    for m in members:
        m.step(steps)
    seq = time.perf_counter() - t0
    with pkg.Ensemble(members) as ens:
        ens.step(2)  # capture + first replay
        t0 = time.perf_counter()
        ens.step(steps)
        par = time.perf_counter() - t0
    tiles = sorted({(c[0], m.tuned_tile()) for c, m in zip(cfgs, members)})
    cells = sum(int(np.prod(n)) for n in shapes)
    for m in members:
        m.close()
    return {
        "metric": f"ensemble solves/s ({len(shapes)} Landau solves of {steps} steps, uneven grids, FP32 f)",
        "value": len(shapes) / par, "unit": "solves/s", "n_gpus": 1, "steps": steps, "warmup": 2,
        "ms_per_step": 1e3 * par / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-storage, f32 interpolation arithmetic",
        "data": f"synthetic: Landau initial conditions, alpha ~ U[0.01, 0.5], k in {{0.3, 0.5}}, seed {seed}",
        "config": {"workload": "batched ensemble: 16 x 32^4, 16 x 48^2 x 64^2, 16 x 64^4, one HIP graph per two steps",
                   "members": len(shapes), "cells": cells, "tuned_tiles": [[list(t[0]), list(t[1])] for t in tiles]},
        "cell_updates_per_s": cells * steps / par,
        "sequential_solves_per_s": len(shapes) / seq,
        "speedup_vs_sequential": seq / par,
    }


def adjoint_bench(pkg, steps, n=96, every=10, seed=20261203):
    """Gradient of the integrated field energy J = dt sum_k |E(f^k)|^2 dx dy
    over `steps` steps (vlp_adjoint_field_energy, FP64) from a Landau state
    at n^4 (box 4 pi x 4 pi x [-6, 6]^2, dt 0.02, idcase 10) plus a seeded
    density perturbation M(v) (P delta)(x, y), delta on a 64 x 64 grid, P the
    trigonometric interpolation to n x n: the sweep of the adjoint workload
    test, timed against the same number of plain steps.  Only the velocity
    moment of the gradient (weight M) crosses to the host."""
    import numpy as np
    nb = 2 * steps + 2
    cfg = pkg.VlpConfig.make([n] * 4, [0.0, 0.0, -6.0, -6.0], [4 * math.pi, 4 * math.pi, 6.0, 6.0], 10, 0.02, nb)
    rng = np.random.default_rng(seed)
    delta = rng.normal(0.0, 1e-3, (64, 64))
    P = pkg.trig_interp_matrix(64, n)
    v = -6.0 + np.arange(n) * (12.0 / n)
    M = np.exp(-0.5 * (v[:, None] ** 2 + v[None, :] ** 2)) / (2 * math.pi)
    with pkg.Vlp4d(cfg) as sim:
        sim.set_f(sim.get_f() + M[:, :, None, None] * (P @ delta @ P.T)[None, None, :, :])
        sim.field_energy_gradient(1, vweight=M, checkpoint_every=1)  # warm-up: operators, buffers
        t0 = time.perf_counter()
        J, _ = sim.field_energy_gradient(steps, vweight=M, checkpoint_every=every)
        t_adj = time.perf_counter() - t0
        t0 = time.perf_counter()
        sim.step(steps)
        t_fwd = time.perf_counter() - t0
    return {
        "metric": f"adjoint field-energy sweep ({steps} steps at {n}^4, FP64, checkpoint every {every})",
        "value": t_adj, "unit": "s", "n_gpus": 1, "steps": steps, "warmup": 1, "ms_per_step": 1e3 * t_adj / steps,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": f"synthetic: Landau state + density perturbation on a 64 x 64 grid, seed {seed}",
        "config": {"workload": f"vlp_adjoint_field_energy over {steps} steps, {n}^4, velocity moment of the gradient",
                   "grid": [n] * 4, "checkpoint_every": every},
        "J": J, "forward_s": t_fwd, "sweep_over_forward": t_adj / t_fwd,
    }


def lagrange_cpu_baseline(n, steps=1):
    """The Lagrange oracle (C/OpenMP port of miniapps/vlp4d/openmp) on the
    host cores: `steps` onetimestep(s) of the same deck, timed in-process."""
    sys.path.insert(0, ROOT)
    import oracle
    cores = min(16, os.cpu_count() or 1)
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    oracle.load()
    cfg = oracle.make_config([n] * 4, SLD10["minphy"], SLD10["maxphy"], SLD10["idcase"], SLD10["dt"], steps)
    st = oracle.LagState(cfg)
    t0 = time.perf_counter()
    st.step(steps)
    sec = (time.perf_counter() - t0) / steps
    st.close()
    return {"value": n ** 4 / sec, "unit": "cell-updates/s", "cores": int(os.environ["OMP_NUM_THREADS"]),
            "kind": "port", "sample": f"{steps} onetimestep(s) of the SLD10 deck at {n}^4, oracle/vlp4d_lag.c "
                                      f"(FP64, OpenMP, {sec:.2f} s/step)"}


def lagrange_bench(pkg, n, steps, warmup, with_cpu):
    """The Lagrange split-step solver (miniapps/vlp4d) on the SLD10 deck at
    n^4: six 1-D order-5 advections per step (x/2, y/2, vx, vy, x/2, y/2),
    each streaming f once (16 B per cell in FP64), plus the field solve; the
    x/2 + y/2 pairs run as one fused pass where the (x, y) plane fits in LDS."""
    nb = warmup + steps
    cfg = pkg.VlpConfig.make(n, SLD10["minphy"], SLD10["maxphy"], SLD10["idcase"], SLD10["dt"], nb)
    sim = pkg.Vlp4d(cfg, scheme=pkg.SCHEME_LAGRANGE)
    if warmup:
        sim.step(warmup)
    tb = sim.timers()
    t0 = time.perf_counter()
    sim.step(steps)
    dt = time.perf_counter() - t0
    ta = sim.timers()
    d = sim.diags()
    fused = bool(sim.kernel_variants() & 1)
    sim.close()
    ph = {lab: (ta[lab][0] - tb[lab][0]) / steps for lab in ("advec2D", "field", "advec4D")}
    # the x/y phase holds 2 launches of the fused pass (kernel_variants bit 0)
    # or 2 x (k_lag_x + k_lag_y); the velocity phase k_lag_v<2> + k_lag_v<3>.
    # The dominant family is the one with the larger share of the step.
    xy = ("k_lag_xy", ph["advec2D"] / 2) if fused else ("k_lag_x/k_lag_y", ph["advec2D"] / 4)
    kern, launch_s = max([xy + (ph["advec2D"],), ("k_lag_v", ph["advec4D"] / 2, ph["advec4D"])],
                         key=lambda r: r[2])[:2]
    launch_bytes = 16 * n ** 4
    achieved = launch_bytes / launch_s / 1e9
    res = {
        "metric": f"4D cell updates/s, Lagrange split-step vlp4d at {n}^4 + % of HBM roofline",
        "value": n ** 4 * steps / dt, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * dt / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic: SLD10 Landau-damping initial condition generated on the GPU",
        "config": {"workload": f"vlp4d Lagrange split step (miniapps/vlp4d), SLD10 deck at {n}^4",
                   "grid": [n] * 4, "idcase": 10, "dt": SLD10["dt"]},
        "roofline": {"bound": "hbm", "kernel": kern, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kern, 64, n, LAG_TRAFFIC_FILE),
                     "algorithmic_bytes_per_launch": launch_bytes, "avg_launch_ms": 1e3 * launch_s},
        "step_hbm_fraction": (4 if fused else 6) * launch_bytes / (dt / steps) / 1e9 / HBM_PEAK_GBS,
        "phases_ms": {k: round(1e3 * v, 4) for k, v in ph.items()},
        "final_nrj": float(d["nrj"][-1]),
    }
    if with_cpu:
        res["cpu_baseline"] = lagrange_cpu_baseline(n)
    return res


def run_mode(pkg, n, precision, steps, warmup):
    """Warm-up then `steps` timed vlp_step calls of the SLD10 deck at n^4 in
    one precision mode; the dominant kernel's roofline from the library's
    HIP-event timers (recorded on its own stream around that launch)."""
    nb = warmup + steps
    cfg = pkg.VlpConfig.make(n, SLD10["minphy"], SLD10["maxphy"], SLD10["idcase"], SLD10["dt"], nb)
    prec = {64: pkg.FP64, 32: pkg.FP32, 33: pkg.FP32_ACC}[precision]
    b = 8 if precision == 64 else 4
    sim = pkg.Vlp4d(cfg, precision=prec)
    if warmup:
        sim.step(warmup)
    t_before = sim.timers()
    t0 = time.perf_counter()
    sim.step(steps)
    dt = time.perf_counter() - t0
    t_after = sim.timers()
    phases = {lab: round(1e3 * (t_after[lab][0] - t_before[lab][0]) / max(1, steps), 4)
              for lab in ("splinecoeff_xy", "advec2D", "field", "splinecoeff_vxvy", "advec4D")}
    adv_s = (t_after["advec4D"][0] - t_before["advec4D"][0]) / steps
    variants = sim.kernel_variants()
    adv_kernel = "k_advect4d_ys" if variants & 128 else "k_advect4d_blk"
    adv_bytes = advect4d_bytes(n, b)
    hbm = adv_bytes / adv_s / 1e9
    valu = ADVECT4D_FLOP_PER_CELL * n ** 4 / adv_s / 1e12
    arith32 = precision == 33
    peak = VALU_PEAK_FP32_TFS if arith32 else VALU_PEAK_TFS
    d = sim.diags()
    sim.close()
    step_bytes = algorithmic_bytes(n, b)
    return {
        "value": n ** 4 * steps / dt,
        "ms_per_step": 1e3 * dt / steps,
        "dtype": {64: "f64", 32: "f32-storage, f64 arithmetic", 33: "f32-storage, f32 interpolation arithmetic"}[
            precision],
        # the 4-D advection binds on VALU issue (~860 flop per cell against
        # 60 / 30 algorithmic bytes), so the roofline is the vector-ALU one
        # with the HBM fraction of the same launch beside it
        "roofline": {"bound": "valu", "kernel": adv_kernel, "achieved": valu, "peak": peak, "unit": "TFLOP/s",
                     "frac": valu / peak, "flop_per_cell": ADVECT4D_FLOP_PER_CELL, "avg_launch_ms": 1e3 * adv_s,
                     "peak_note": ("published packed-FP32 vector peak; this build disables packed FP32, whose "
                                   "non-packed rate is 78.6 TFLOP/s, the frac against that is "
                                   f"{valu / VALU_PEAK_TFS:.3f}") if arith32 else "FP64 vector peak",
                     "hbm": {"achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm / HBM_PEAK_GBS,
                             "algorithmic_bytes_per_launch": adv_bytes,
                             "traffic": measured_traffic(adv_kernel, precision, n),
                             "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT)}},
        "step_hbm_fraction": step_bytes / (dt / steps) / 1e9 / HBM_PEAK_GBS,
        "step_algorithmic_bytes": step_bytes,
        "phases_ms": phases,
        "final_nrj": float(d["nrj"][-1]),
        "nrj": [float(x) for x in d["nrj"]],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", type=int, default=0, choices=[0, 64, 32, 33],
                    help="one mode only -- 64: FP64; 32: FP32 storage, FP64 interpolation arithmetic; "
                         "33: FP32 storage and FP32 interpolation arithmetic.  Default (0): 33 (the value), "
                         "64 and 32 in the same run")
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--ensemble", action="store_true",
                    help="batched-ensemble workload (48 uneven Landau solves); --steps is the solve length")
    ap.add_argument("--adjoint", action="store_true",
                    help="the adjoint field-energy sweep at 96^4 (FP64); --steps is the sweep length")
    ap.add_argument("--lagrange", action="store_true",
                    help="the Lagrange split-step solver of miniapps/vlp4d on the same deck")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "the multi-GPU decomposition is a design note, not an implemented path.", file=sys.stderr)
        sys.exit(2)
    if a.calibrate_copy:
        calibrate_copy()
        return
    pkg = load_pkg()
    if a.ensemble:
        print(json.dumps(ensemble_bench(pkg, a.steps)))
        return
    if a.adjoint:
        print(json.dumps(adjoint_bench(pkg, a.steps)))
        return
    if a.lagrange:
        print(json.dumps(lagrange_bench(pkg, a.n, a.steps, a.warmup, not a.no_cpu_baseline)))
        return
    n = a.n
    modes = [a.precision] if a.precision else [33, 64, 32]
    runs = {p: run_mode(pkg, n, p, a.steps, a.warmup) for p in modes}
    head = runs[modes[0]]
    res = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        # the only published figure is the A100 stdpar FP64 MainLoop of this deck
        # (docs/vlp4d.md:148,168-171); no FP32 figure exists in the reference
        "vs_baseline": head["value"] / A100_REF if n == 128 else None,
        "dtype": head["dtype"],
        "data": "synthetic: SLD10 Landau-damping initial condition generated on the GPU",
        "config": {"workload": f"vlp4d spline step, SLD10 deck at {n}^4 (box 12.56637112, v in [-9,9], dt 0.01)",
                   "grid": [n, n, n, n], "idcase": 10, "dt": 0.01,
                   "storage_bits": 64 if modes[0] == 64 else 32},
        "roofline": head["roofline"],
        "step_hbm_fraction": head["step_hbm_fraction"],
        "step_algorithmic_bytes": head["step_algorithmic_bytes"],
        "phases_ms": head["phases_ms"],
        "final_nrj": head["final_nrj"],
    }
    if 64 in runs and modes[0] != 64:
        f64 = runs[64]
        res["fp64"] = {k: f64[k] for k in ("value", "ms_per_step", "dtype", "roofline", "step_hbm_fraction",
                                           "phases_ms", "final_nrj")}
        res["fp64"]["vs_baseline"] = f64["value"] / A100_REF if n == 128 else None
        res["speedup_vs_fp64"] = head["value"] / f64["value"]
        w64 = [math.exp(2 * x) for x in f64["nrj"]]
        for p in runs:
            if p == 64:
                continue
            w = [math.exp(2 * x) for x in runs[p]["nrj"]]
            key = "fp32_acc" if p == 33 else "fp32"
            res.setdefault("field_energy_rel_err_vs_fp64", {})[key] = (
                max(abs(u - v) for u, v in zip(w, w64)) / max(abs(v) for v in w64))
        if 32 in runs:
            r32 = runs[32]
            res["fp32_f64arith"] = {"value": r32["value"], "ms_per_step": r32["ms_per_step"], "dtype": r32["dtype"],
                                    "speedup_vs_fp64": r32["value"] / f64["value"]}
    if not a.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(n)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
