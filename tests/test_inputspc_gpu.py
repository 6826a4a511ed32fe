"""The reference's regression inputs run end to end from their input.spc
(espec_amd/inputspc.py): initial state, LC/TF rule, propagation, C(t)
assembly and spectrum on the device, against the published tables.

Bars: the printed digits (7 decimals: one unit is 1e-7) for the FP64
propagators; ABSTOL = 1e-6 bisection runs (test4 .PSIL, test14 .PPSIL)
move C(t) by a few 1e-6, as in test_dropin_gpu.py.
"""
import os
import shutil

import numpy as np
import pytest

import setups as S
from espec_amd import inputspc as I

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = os.path.join(HERE, "golden", "inputs")


def run(t, base=None):
    print(f"case: input.test{t}", flush=True)
    s = I.read(os.path.join(base or INPUTS, f"input.test{t}"))
    return s, I.run(s)


def test_input_test12_pfspo(gpu):
    s, r = run(12)
    tab, cor, spec = S.load("test12_table"), S.load("test12_corr"), S.load("test12_spec")
    assert r["lc"] == 1220 and abs(r["tf"] - 249.856) < 1e-9
    k = len(tab)
    assert np.abs(r["cr"][:k] - tab[:, 2]).max() <= 1.01e-7
    assert np.abs(r["ci"][:k] - tab[:, 3]).max() <= 1.01e-7
    assert np.abs(r["u"] - cor[:, 1]).max() <= 1.01e-7 and np.abs(r["v"] - cor[:, 2]).max() <= 1.01e-7
    assert np.abs(r["t"] - cor[:, 0]).max() <= 1e-6
    d = [np.abs(r[c] - spec[:, j]).max() for c, j in (("e", 0), ("w", 1), ("wi", 3))]
    print(f"spectrum |dE| {d[0]:.1e} |dW| {d[1]:.1e} |dWi| {d[2]:.1e}", flush=True)
    assert d[0] <= 6e-8 and d[1] <= 1.01e-7 and d[2] <= 1.01e-7


def test_input_test8_onlyspec(gpu, tmp_path):
    shutil.copy(os.path.join(INPUTS, "input.test8"), tmp_path)
    cor = S.load("correl8_corr")
    with open(tmp_path / "correl8.inp", "w") as f:
        f.write("Correlation Function:\n")
        for row in cor:
            f.write(f"{row[0]:15.6f} {row[1]:12.7f} {row[2]:12.7f}\n")
    s, r = run(8, str(tmp_path))
    spec = S.load("test8_spec")
    assert np.abs(r["e"] - spec[:, 0]).max() <= 6e-8
    assert np.abs(r["w"] - spec[:, 1]).max() <= 6e-8
    assert np.abs(r["wi"] - spec[:, 3]).max() <= 6e-8


def test_input_test4_psil_file_hwhm(gpu):
    s, r = run(4)
    tab, cor, spec = S.load("test4_table"), S.load("test4_corr"), S.load("test4_spec")
    k = len(tab)
    assert np.abs(r["cr"][:k] - tab[:, 2]).max() <= 5e-6
    assert np.abs(r["u"] - cor[:, 1]).max() <= 5e-6 and np.abs(r["v"] - cor[:, 2]).max() <= 5e-6
    d = [np.abs(r[c] - spec[:, j]).max() for c, j in (("e", 0), ("w", 1), ("wi", 3))]
    print(f"spectrum |dE| {d[0]:.1e} |dW| {d[1]:.1e} |dWi| {d[2]:.1e}", flush=True)
    assert d[0] <= 6e-8 and d[1] <= 5e-6 and d[2] <= 5e-6


def test_input_test14_ppsil_lanczs(gpu):
    """Initial state from LANCZS on the device, PPLNZ without a laser.

    The table prints C(t) with 5 decimals (F8.5), so the unrounded C(t)
    here differs from a printed value by up to half a unit, 5e-6, on top
    of any real difference.  The real difference is the drift of the
    ABSTOL = 1e-6 bisection eigenvalues (dstebz.f): each step's Krylov
    eigenvalues are interval midpoints, which rounding-level differences
    between builds move by up to ABSTOL.  Against the 7-decimal tables of
    the same kind of run (test4 at order 10, test13 at order 15, 124928
    steps) that drift measures 3.9e-6 and 5.4e-6; the allowance here is
    about three times the larger, 1.5e-5.  Bar: 5e-6 + 1.5e-5 = 2.0e-5.
    The LANCZS initial state itself is checked against the dense-eigh
    ground state, so it contributes nothing: for unit states a, b and a
    unitary step U, |<a|Ua> - <b|Ub>| <= 2 |a - b|."""
    s, r = run(14)
    tab = S.load("test14_table")
    k = len(tab)
    d = max(np.abs(r["cr"][:k] - tab[:, 3]).max(), np.abs(r["ci"][:k] - tab[:, 4]).max())
    exact = S.test_case(14)["u"]
    u0 = r["psi0"][0]
    dist = min(np.linalg.norm(u0 - exact), np.linalg.norm(u0 + exact))
    print(f"|dC| {d:.2e}  |psi0 - psi_exact| {dist:.2e}", flush=True)
    assert dist <= 1e-10
    assert d <= 5e-6 + 1.5e-5


@pytest.mark.parametrize("t,method,bar", [(5, ".PFSOD", 2.01e-7), (6, ".PFSPO", 2.01e-7)])
def test_input_morse_fft_propagators(gpu, t, method, bar):
    """Morse surfaces with the initial state from LANCZS on the device
    (ABSTOL 1e-6, 13 states for test5): C(t) to two units of the printed
    7th decimal over the published rows (test6's first 1200)."""
    s, r = run(t)
    assert s.propag == method
    tab = S.load(f"test{t}_table")
    k = len(tab)
    d = max(np.abs(r["cr"][:k] - tab[:, 2]).max(), np.abs(r["ci"][:k] - tab[:, 3]).max())
    print(f"|dC| {d:.1e} over {k} rows", flush=True)
    assert d <= bar
    if t == 5:
        spec = S.load("test5_spec")
        dw = max(np.abs(r["w"] - spec[:, 1]).max(), np.abs(r["wi"] - spec[:, 3]).max())
        print(f"spectrum |dW| {dw:.1e}", flush=True)
        assert np.abs(r["e"] - spec[:, 0]).max() <= 6e-8 and dw <= bar


def test_input_test13_psil_order15(gpu):
    """.PSIL at Krylov order 15 from the LANCZS ground state, ABSTOL 1e-6
    bisection in both.  Each step's Krylov eigenvalues sit anywhere in
    their 1e-6 bisection interval, so over 124928 steps C(t) drifts by a
    few 1e-6 from the reference's (5.4e-6 at order 15; test4 at order 10
    reaches 3.9e-6): the bar is 1e-5, the spectrum's 5e-6."""
    s, r = run(13)
    assert (s.propag, s.mp) == (".PSIL", 15)
    tab, cor, spec = S.load("test13_table"), S.load("test13_corr"), S.load("test13_spec")
    k = len(tab)
    d = max(np.abs(r["cr"][:k] - tab[:, 2]).max(), np.abs(r["ci"][:k] - tab[:, 3]).max())
    dw = max(np.abs(r["w"] - spec[:, 1]).max(), np.abs(r["wi"] - spec[:, 3]).max())
    print(f"|dC| {d:.1e}  spectrum |dW| {dw:.1e}", flush=True)
    assert d <= 1e-5 and np.abs(r["u"] - cor[:, 1]).max() <= 1e-5 and dw <= 5e-6
