"""Host-side checks that need no GPU: deck parsing (A1), the C-ABI surface,
configuration errors, the command-line drivers' exit paths and bench.py's
one-GPU guard.

Deck values are pinned by what the reference's own decks literally contain
(wk/SLD10.dat, wk/TSI20.dat, copied under tests/golden/decks).
"""
import ctypes
import math
import os
import re
import signal
import subprocess
import sys

import numpy as np
import pytest

from conftest import DECKS, ROOT, sld_box

DECK_FILES = ["SLD10.dat", "TSI20.dat", "SLD10_large.dat", "prof_SLD10.dat"]


def test_parse_sld10(orc, vlp):
    for parse in (orc.parse_deck, vlp.parse_deck):
        c = parse(os.path.join(DECKS, "SLD10.dat"))
        assert list(c.nxmax) == [128, 128, 128, 128]
        assert list(c.minPhy) == [0.0, 0.0, -9.0, -9.0]
        assert list(c.maxPhy) == [12.56637112, 12.56637112, 9.0, 9.0]
        assert (c.idcase, c.dt, c.nbiter, c.ifreq, c.fxvx) == (10, 0.01, 40, 8, 0)
        assert c.dx[0] == 12.56637112 / 128 and c.dx[2] == 18.0 / 128


def test_parse_tsi20(orc, vlp):
    for parse in (orc.parse_deck, vlp.parse_deck):
        c = parse(os.path.join(DECKS, "TSI20.dat"))
        assert list(c.nxmax) == [512, 4, 512, 4]
        assert c.maxPhy[0] == 25.13274122871834590
        assert (c.idcase, c.dt, c.nbiter, c.ifreq, c.fxvx) == (20, 0.125, 512, 8, 1)


@pytest.mark.parametrize("deck", DECK_FILES)
def test_product_parser_equals_oracle(orc, vlp, deck):
    a = orc.parse_deck(os.path.join(DECKS, deck))
    b = vlp.parse_deck(os.path.join(DECKS, deck))
    for f in ("nxmax", "minPhy", "maxPhy", "dx"):
        assert list(getattr(a, f)) == list(getattr(b, f))
    for f in ("idcase", "dt", "nbiter", "ifreq", "fxvx"):
        assert getattr(a, f) == getattr(b, f)


def test_idcase_read_as_six_characters(vlp, tmp_path):
    # fgets(idcase, 7, ...) then atoi (VMO/Init.cpp:53-56,78): only the first
    # six characters after the colon count
    txt = open(os.path.join(DECKS, "SLD10.dat")).read()
    p = tmp_path / "d.dat"
    for spaces, expect in ((4, 20), (5, 2), (6, 0)):
        p.write_text(txt.replace("test case : 10", "test case :" + " " * spaces + "20"))
        assert vlp.parse_deck(str(p)).idcase == expect


def test_missing_deck(vlp, orc, tmp_path):
    with pytest.raises(vlp.Vlp4dError) as ei:
        vlp.parse_deck(str(tmp_path / "nope.dat"))
    assert ei.value.code == vlp.E_FILE
    with pytest.raises(RuntimeError):
        orc.parse_deck(str(tmp_path / "nope.dat"))


def test_header_declares_exactly_the_exports(vlp):
    hdr = open(os.path.join(ROOT, "include", "vlp4d.h")).read()
    declared = sorted(set(re.findall(r"\b(vlp_[a-z0-9_]+)\s*\(", hdr)))
    assert declared == sorted(vlp.EXPORTS)
    lib = vlp.load()
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.vlp_version().startswith(b"vlp4d-mi355x")


def test_timer_labels_match_reference_order(vlp):
    lib = vlp.load()
    got = [lib.vlp_timer_label(k).decode() for k in range(13)]
    assert got == ["total", "MainLoop", "pack", "comm", "unpack", "advec2D", "advec4D", "field",
                   "all_reduce", "Fourier", "diag", "splinecoeff_xy", "splinecoeff_vxvy"]


@pytest.mark.parametrize("mutate,code", [
    (lambda c: c.nxmax.__setitem__(0, 8), "E_SIZE"),      # N < MMAX+2
    (lambda c: c.nxmax.__setitem__(2, 260), "E_SIZE"),    # line buffer
    (lambda c: setattr(c, "idcase", 11), "E_CASE"),       # advertised, not implemented
    (lambda c: setattr(c, "dt", -0.1), "E_ARG"),
])
def test_config_errors_before_any_device_call(vlp, mutate, code):
    c = vlp.VlpConfig.make(*sld_box(16))
    mutate(c)
    with pytest.raises(vlp.Vlp4dError) as ei:
        vlp.Vlp4d(c)
    assert ei.value.code == getattr(vlp, code)


def test_create_without_device_is_an_error_not_a_crash(vlp):
    c = vlp.VlpConfig.make(*sld_box(16))
    try:
        with vlp.Vlp4d(c) as sim:
            assert sim.iter == 0  # a GPU is present
    except vlp.Vlp4dError as e:
        assert e.code == vlp.E_HIP


def _run(args, cwd):
    return subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("exe", ["p3-miniapps_amd/vlp4d", "oracle/vlp4d_oracle"])
def test_driver_missing_deck_aborts_like_reference(exe, tmp_path):
    p = _run([os.path.join(ROOT, exe), "-f", "absent.dat"], tmp_path)
    assert "import: error file not found" in p.stdout
    assert p.returncode == -signal.SIGABRT


@pytest.mark.parametrize("exe", ["p3-miniapps_amd/vlp4d", "oracle/vlp4d_oracle"])
def test_driver_unknown_case_aborts_like_reference(exe, tmp_path):
    txt = open(os.path.join(DECKS, "SLD10.dat")).read().replace("test case : 10", "test case : 11")
    (tmp_path / "d.dat").write_text(txt.replace(": 128", ": 12"))
    p = _run([os.path.join(ROOT, exe), "-f", "d.dat"], tmp_path)
    assert "Unknown test case !" in p.stdout
    assert p.returncode == -signal.SIGABRT


def test_oracle_driver_outputs(tmp_path):
    txt = open(os.path.join(DECKS, "SLD10.dat")).read().replace(": 128", ": 12")
    txt = txt.replace("Time step : 0.01", "Time step : 0.05").replace("iterations : 40", "iterations : 3")
    (tmp_path / "d.dat").write_text(txt)
    p = _run([os.path.join(ROOT, "oracle", "vlp4d_oracle"), "-f", "d.dat"], tmp_path)
    assert p.returncode == 0, p.stderr
    assert [ln for ln in p.stdout.splitlines() if ln.startswith("iter ")] == ["iter 0", "iter 1", "iter 2"]
    labels = [ln.split()[0] for ln in p.stdout.splitlines() if " [s], " in ln]
    assert labels[:3] == ["total", "MainLoop", "pack"] and len(labels) == 13
    nrj = (tmp_path / "nrj.out").read_text().splitlines()
    l2 = (tmp_path / "l2norm.out").read_text().splitlines()
    assert len(nrj) == 4 and len(l2) == 4
    # %17.13e columns (VMO/Diags.cpp:408,418)
    assert re.fullmatch(r"(-?\d\.\d{13}e[+-]\d\d\s?){5}", nrj[1].strip() + " ")
    assert float(nrj[2].split()[0]) == 2 * 0.05


def test_bench_refuses_more_than_one_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2 and "one GPU" in p.stderr


def test_committed_traffic_summaries_parse():
    """bench.py reads each kernel's measured HBM bytes per launch from the
    committed rocprofv3 summaries; every entry parses and lies between the
    kernel's algorithmic bytes (a lower bound on what it must move) and
    four times that."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    n = 128
    esz = {64: 8, 33: 4, 32: 4}
    for prec, b in esz.items():
        t = bench.measured_traffic("k_advect4d_ys", prec, n)
        alg = bench.advect4d_bytes(n, b)
        assert t is not None and alg <= t <= 4 * alg, (prec, t, alg)
    # the Lagrange passes each stream f once: 16 bytes per cell in FP64
    for kern in ("k_lag_xy", "k_lag_v"):
        t = bench.measured_traffic(kern, 64, n, bench.LAG_TRAFFIC_FILE)
        assert t is not None and 16 * n ** 4 <= t <= 4 * 16 * n ** 4, (kern, t)
    assert bench.measured_traffic("k_advect4d_ys", 64, 96) is None
