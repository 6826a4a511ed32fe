"""The reference's own main program (espec.f, compiled from its sources)
with the hot path linked to the drop-in (oracle/_ref/espec_dropin.x, built
by `make -C oracle dropin`), run on the GPU on the reference's regression
inputs: its printed C(t) table, correlation function and spectrum against
the published results (tests/golden, extracted by make_fixtures.py).

Everything before and after the propagation (input reader, potentials,
the file handling of espec.f) is the reference's own code; the initial
eigenstates (MTRXDIAG, LANCZS), the propagator, C(t) assembly pieces,
HWHM and SPECTRUMTD are the drop-in's.

The reference is the same v0.7 program built entirely from the reference's
sources (tests/golden/v07_*, written by make_ref_fixtures.py from
oracle/_ref/espec_ref.x): the published v0.4 results differ from v0.7 by
the potential shift it applies (3.3e-2 in test12's C(t)), which test4's
.FILE potential does not have, so test4 is also held to the published
table.  Bars: one unit of the printed 7th decimal for the FFT propagator
(the drop-in's REF mode matches the reference build to 1e-12, so only
rounding flips of the printed digit remain), a few 1e-6 for the
ABSTOL = 1e-6 Lanczos runs (bisection midpoints), one unit of F8.5 for
PPLNZ's table."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import make_fixtures as MF
import setups as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "espec_dropin.x")
INPUTS = os.path.join(ROOT, "tests", "golden", "inputs")
STOPS = ["Correlation Function", "Spectrum:", "Partial auto", "Total correlation",
         "Calculating", "Propagating", "Eigenvector propagated"]


def run_main(t, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/espec_dropin.x not built (make -C oracle dropin)")
    shutil.copy(os.path.join(INPUTS, f"input.test{t}"), tmp_path / "input.spc")
    for f in os.listdir(INPUTS):
        if not f.startswith("input."):
            shutil.copy(os.path.join(INPUTS, f), tmp_path / f)
    print(f"case: espec_dropin.x input.test{t}", flush=True)
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=600)
    out = r.stdout.splitlines()
    assert "Propagation operator error" not in r.stdout, r.stdout[-3000:]
    return out


def arr(rows):
    return np.array([[float(x.replace("D", "E")) for x in r.split()] for r in rows])


def blocks(lines, table_cols=6, table_title="Partial auto-correlation function"):
    tab = arr(MF.block(lines, table_title, table_cols,
                       [s for s in STOPS if not table_title.startswith(s)]))
    cor = arr(MF.block(lines, "Correlation Function", 3, [s for s in STOPS if s != "Correlation Function"]))
    spec = arr(MF.block(lines, "Spectrum:", 4, [s for s in STOPS if s != "Spectrum:"]))
    return tab, cor, spec


def test_main_test12_pfspo(gpu, tmp_path):
    """.PFSPO, 2.5e6 steps of a 256-point packet, MTRXDIAG initial state."""
    tab, cor, spec = blocks(run_main(12, tmp_path))
    gt, gc, gs = S.load("v07_test12_table"), S.load("v07_test12_corr"), S.load("v07_test12_spec")
    assert tab.shape == gt.shape and cor.shape == gc.shape and spec.shape == gs.shape
    d = [np.abs(tab[:, 2:4] - gt[:, 2:4]).max(), np.abs(cor[:, 1:] - gc[:, 1:]).max(),
         np.abs(spec[:, 1:] - gs[:, 1:]).max()]
    print(f"test12 |dC| table {d[0]:.1e} corr {d[1]:.1e} spectrum {d[2]:.1e}", flush=True)
    assert max(d) <= 1.01e-7
    assert np.abs(spec[:, 0] - gs[:, 0]).max() <= 6e-8
    assert np.abs(tab[:, 1] / gt[:, 1] - 1).max() <= 2e-6  # E12.6


def test_main_test4_psil(gpu, tmp_path):
    """.PSIL order 10 from a .FILE potential and .GETR state, ABSTOL 1e-6.

    The table against the v0.7 build and the published v0.4 results; the
    assembled correlation function against the published one only: the
    v0.7 build prints Re C = 0 at 30 of its 4096 rows (1.2 <= |t| <= 3.6
    fs), where the v0.4 results and the drop-in both print values up to
    0.37 in size that continue their neighbours smoothly, so those rows are
    an artefact of that build."""
    tab, cor, spec = blocks(run_main(4, tmp_path))
    for pre in ("v07_", ""):
        gt = S.load(pre + "test4_table")
        assert tab.shape == gt.shape
        d = np.abs(tab[:, 2:4] - gt[:, 2:4]).max()
        print(f"test4 table against {pre or 'v0.4 '}|dC| {d:.1e}", flush=True)
        assert d <= 6e-6
    gc = S.load("test4_corr")
    assert cor.shape == gc.shape
    d = np.abs(cor[:, 1:] - gc[:, 1:]).max()
    print(f"test4 correlation function |dC| {d:.1e}", flush=True)
    assert d <= 6e-6


def test_main_test14_ppsil(gpu, tmp_path):
    """.PPSIL (PPLNZ) from the LANCZS ground state; PPLNZ prints its own
    7-column table (t, ET, E0, Pr, Pi, |P0|^2, norm) with F8.5 C(t)."""
    lines = run_main(14, tmp_path)
    tab = arr(MF.block(lines, "Auto-correlation function", 7, ["Eigenvector propagated"]))
    gt = S.load("v07_test14_table")
    assert tab.shape == gt.shape
    d = np.abs(tab[:, 3:5] - gt[:, 3:5]).max()
    print(f"test14 |dP| {d:.1e}", flush=True)
    assert d <= 1.01e-5  # one unit of F8.5's last digit (bisection drift)
