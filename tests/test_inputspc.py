"""input.spc reader (espec_amd/inputspc.py) on the reference's own test
inputs (copies under tests/golden/inputs): every keyword the propagation
path reads, and the grid, masses and potentials built from them against
the hand-made setups of tests/golden/setups.py, which the published tables
pin (tests/test_oracle_golden.py)."""
import os

import numpy as np
import pytest

import setups as S
from espec_amd import inputspc as I

HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = os.path.join(HERE, "golden", "inputs")


def spc(t):
    return I.read(os.path.join(INPUTS, f"input.test{t}"))


def test_keywords_test12():
    s = spc(12)
    assert (s.tpcalc, s.tdi, s.dim, s.np) == (".SPECTRUM", ".TD", ".1D", [256])
    assert s.potential == ".OHS" and s.cst_i == [1.146, 1943.79] and s.cst_f == [1.211, 1629.23]
    assert (s.xi, s.xf, s.mass) == ([0.5], [2.8], [7.49])
    assert (s.inieigvc, s.change, s.prtcrl, s.tpdiag) == (".CALC", True, ".PARTIAL", ".MTRXDIAG")
    assert (s.nist, s.abstol) == ([7, 0], 1e-6)
    assert (s.propag, s.ti, s.tf, s.dt, s.prpgstate) == (".PFSPO", 0.0, 250.0, 1e-4, 0)
    assert (s.tptrans, s.prptol, s.mf, s.window, s.wp, s.hwhm) == (".ONE", 3.0, 12, ".SG", 1e-3,
                                                                      None)


def test_keywords_other_inputs():
    s = spc(4)
    assert (s.potential, s.pot_file, s.inieigvc) == (".FILE", "eigpot_test3.inp", ".GETR")
    assert (s.propag, s.ti, s.tf, s.dt, s.mp) == (".PSIL", 0.0, 250.0, 2e-3, 10)
    assert s.hwhm == (0.087, 0.0, 0.12, 19.562) and s.window == ".NONE"
    s = spc(5)
    assert s.potential == ".MORSE" and s.cst_i == [1.146, 1943.79, 13.71]
    assert (s.propag, s.dt, s.mf, s.wp, s.nist) == (".PFSOD", 5e-5, 14, 1e-5, [13, 0])
    s = spc(8)
    assert (s.tpcalc, s.corr_file, s.mf, s.hwhm) == (".ONLYSPEC", "correl8.inp", 13,
                                                     (0.08, 0.0, 0.0, 0.0))
    s = spc(14)
    assert (s.propag, s.dt, s.mp, s.np, s.tpdiag) == (".PPSIL", 2e-3, 10, [201], ".LANCZS")
    assert spc(10).np == [250] and spc(10).propag == ".PFSOD"


def test_list_directed_records():
    s = I.parse("**MAIN\n*DIMENSION\n.1D ! one axis\n 64, /\n*MASS\n 1.5D0 2/\n"
                "**TD\n*PROPAG\n.PSIL   ! Lanczos\n 0.0  1.0D+1  2.5d-3  14 /\n**END\n")
    assert s.np == [64] and s.mass == [1.5, 2.0]
    assert (s.propag, s.ti, s.tf, s.dt, s.mp) == (".PSIL", 0.0, 10.0, 2.5e-3, 14)


@pytest.mark.parametrize("t", [12, 10, 5, 14])
def test_built_inputs_match_setups(t):
    s = spc(t)
    x, sh = I.grid(s)
    c = S.test_case(t)
    assert np.allclose(sh, c["sh"][0], rtol=1e-15, atol=0)
    m = s.mass[0] * I.AMU
    assert m == c["sm"][0]
    v1 = I.potential(s.potential, s.cst_f, x, m)
    assert np.allclose(v1, c["vpot"], rtol=1e-13, atol=1e-18)


def test_file_sections_test4():
    s = spc(4)
    psi, pot = I._sections(os.path.join(INPUTS, s.pot_file))[:2]
    c = S.test_case(4)
    assert np.array_equal(psi[:, 1], c["u"]) and np.array_equal(pot[:, 1], c["vpot"])


@pytest.mark.parametrize("block,attr,word", [
    ("*ABSORBING\n.SMOOTHW\n0.5\n1.0 2.0\n", "tpabsor", ".SMOOTHW"),
    ("*ABSORBING\n.VOPTIC_LINEAR\n0.5\n1.0 2.0\n", "tpabsor", ".VOPTIC_LINEAR"),
    ("*PULSE_AND_DIPOLE\n.GAUS\n1e12 10 3.0 20 0\n.NULL\n", "efc", ".GAUS")])
def test_absorber_and_pulse_are_refused_by_run(block, attr, word):
    """*ABSORBING (rdinput.f:526) other than .NULL and *PULSE_AND_DIPOLE
    (rdinput.f:549) other than .NONE/.NULL change the propagation; run()
    refuses them instead of propagating without them."""
    text = open(os.path.join(INPUTS, "input.test12")).read()
    k = text.index("*PROPAG")
    s = I.parse(text[:k] + block + text[k:], INPUTS)
    assert getattr(s, attr) == word
    with pytest.raises(ValueError, match=attr == "tpabsor" and "ABSORBING" or "PULSE"):
        I.run(s)


def test_null_absorber_and_no_pulse_parse():
    text = open(os.path.join(INPUTS, "input.test12")).read()
    k = text.index("*PROPAG")
    s = I.parse(text[:k] + "*ABSORBING\n.NULL\n0\n0 0\n*PULSE_AND_DIPOLE\n.NONE\n0\n.NULL\n"
                + text[k:], INPUTS)
    assert (s.tpabsor, s.efc) == (".NULL", ".NONE")
