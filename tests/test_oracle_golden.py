"""The oracle against the reference's published regression outputs.

Fixtures under tests/golden/ hold the numeric blocks of the reference's
ttests/results.testN (made by tests/golden/make_fixtures.py); inputs are
rebuilt by tests/golden/setups.py from the matching input.testN.

Printed precision sets the bars: C(t) is printed with 7 decimals (F12.7),
so agreement within 6e-8 is agreement to the last printed digit.  Wider
bars, each with its reason:
  test4  (.PSIL, ABSTOL = 1e-6): psi0 and V come from a file printed to 6
         significant digits, and with ABSTOL > 0 each Krylov eigenvalue is a
         bisection midpoint known only to ABSTOL; 2e-6 over 200 samples.
  test10 (.PFSOD, N = 250): the reference's radix-5 FFT twiddles carry its
         single-precision 2*pi (fft3d.f:119); the oracle's non-power-of-two
         transform is an exact DFT, so the phases drift by ~1e-6.
"""
import numpy as np
import pytest

import setups as S
from oracle import EXACT, REF

# test: (samples compared, |dC| bar)
TABLE_CASES = {12: (60, 6e-8), 6: (40, 6e-8), 5: (40, 6e-8), 4: (200, 2e-6), 10: (3, 5e-6)}


def propagate(orc, c, nsamp):
    lc, tf = orc.lc_rule(c["ti"], c["tf"], c["dt"], c["mf"], True)
    tfk = c["ti"] + (nsamp - 1) * lc * c["dt"]
    if c["method"] == "PFSPO":
        o = orc.pspofft(c["np_"], c["dt"], c["ti"], tfk, lc, c["sh"], c["sm"], c["u"], c["v"],
                        c["vpot"], REF, energies=True)
        mag = np.hypot(o["cr"], o["ci"])          # pspofft.f:148 prints |C|
    elif c["method"] == "PFSOD":
        o = orc.psodfft(c["np_"], c["dt"], c["ti"], tfk, lc, c["sh"], c["sm"], c["u"], c["v"],
                        c["vpot"], REF)
        mag = np.hypot(o["cr"], o["ci"])
    else:
        o = orc.plnz(c["np_"], c["mp"], c["abstol"], c["dt"], c["ti"], tfk, lc, c["shm"],
                     c["u"], c["v"], c["vpot"], REF)
        mag = o["cr"] ** 2 + o["ci"] ** 2         # plnz.f:155 prints |C|^2
    return lc, tf, o, mag


@pytest.mark.parametrize("t", sorted(TABLE_CASES))
def test_correlation_table(orc, t):
    nsamp, tol = TABLE_CASES[t]
    c = S.test_case(t)
    g = S.load(f"test{t}_table")
    lc, tf, o, mag = propagate(orc, c, nsamp)
    # the LC/TF rule (espec.f:1330-1350): sample spacing and adjusted TF
    assert abs(g[1, 0] - lc * c["dt"]) < 1e-9
    assert abs(tf - (149.9136 if t == 6 else 249.856)) < 1e-9
    k = min(nsamp, len(o["cr"]))
    assert k == nsamp
    assert np.abs(o["cr"][:k] - g[:k, 2]).max() <= tol
    assert np.abs(o["ci"][:k] - g[:k, 3]).max() <= tol
    assert np.abs(mag[:k] - g[:k, 4]).max() <= tol
    if "e" in o:  # energy column, 6 significant digits
        assert np.abs(o["e"][:k] / g[:k, 1] - 1).max() <= 2e-6


@pytest.mark.parametrize("t,mf", [(12, 12), (10, 12), (5, 14), (4, 12)])
def test_correlation_assembly(orc, t, mf):
    """espec.f:1546-1566: the 2^(MF-1) samples mirrored into C(t) on
    t = -TF + (i-1) DTF; matches the printed 'Correlation Function'."""
    tab, cor = S.load(f"test{t}_table"), S.load(f"test{t}_corr")
    mc = (1 << mf) // 2
    u, v = orc.assemble_corr(mf, tab[:mc, 2], tab[:mc, 3])
    assert np.array_equal(u, cor[:, 1]) and np.array_equal(v, cor[:, 2])
    dtf = tab[1, 0] - tab[0, 0]
    assert np.allclose(cor[:, 0], cor[0, 0] + np.arange(2 * mc) * dtf, rtol=0, atol=6e-4)


# (test, MF, window (1 = .SG), WP, HWHM (eV) or None, DE - E1I, ADE, corr fixture)
SPECTRA = [
    (12, 12, 1, 1e-3, None, 0.0, 0.0, "test12_corr"),
    (10, 12, 1, 1e-3, None, 0.0, 0.0, "test10_corr"),
    (5, 14, 1, 1e-5, None, 0.0, 0.0, "test5_corr"),
    (4, 12, 0, 0.0, 0.087, -0.12, 19.562, "test4_corr"),
    (8, 13, 0, 0.0, 0.08, 0.0, 0.0, "correl8_corr"),
]


@pytest.mark.parametrize("mode", [REF, EXACT])
@pytest.mark.parametrize("t,mf,iwin,wp,hw,de,ade,src", SPECTRA)
def test_spectrum(orc, t, mf, iwin, wp, hw, de, ade, src, mode):
    """HWHM (hwhm.f) then SPECTRUMTD (spectrumtd.f: window, half swap,
    FFTNT, printed E, W = FFTr, FFTi) against the printed 'Spectrum'."""
    cor, spec = S.load(src), S.load(f"test{t}_spec")
    tf = -cor[0, 0]
    dtf = tf / (1 << (mf - 1))          # DTF = TF/MC (espec.f:1600-1601)
    u, v = cor[:, 1].copy(), cor[:, 2].copy()
    if hw is not None:
        u, v = orc.hwhm(mf, dtf, tf, hw, de, u, v)
    e, w, wi = orc.spectrumtd(iwin, mf, dtf, tf, wp, ade, u, v, mode)
    assert np.abs(e - spec[:, 0]).max() <= 6e-8
    assert np.abs(w - spec[:, 1]).max() <= 6e-8
    assert np.abs(w - spec[:, 2]).max() <= 6e-8
    assert np.abs(wi - spec[:, 3]).max() <= 6e-8


def test_test8_input_is_printed_correlation():
    """The .ONLYSPEC input file and the printed correlation are the same
    numbers (getcorr.f reads what espec.f:1605-1612 prints back)."""
    assert np.array_equal(S.load("correl8_corr"), S.load("test8_corr"))


# LANCZS eigenvalues printed by the reference ('state has been calculating',
# lanczs.f:1031-1032, 8 significant digits) with ABSTOL = 1e-6: the
# bisection leaves each value anywhere in its final interval, so the bar is
# ABSTOL / 2 plus the last printed digit.
LANCZS_GOLD = {
    14: ("ohs", 201, [0.44281536E-02, 0.13282522E-01, 0.22133939E-01, 0.30979451E-01,
                      0.39816698E-01, 0.48643908E-01, 0.57458719E-01]),
    5: ("morse", 256, [0.44125657E-02, 0.13143364E-01, 0.21748722E-01, 0.30226180E-01,
                       0.38576558E-01, 0.46799856E-01, 0.54894435E-01, 0.62860294E-01]),
}


@pytest.mark.parametrize("t", sorted(LANCZS_GOLD))
def test_lanczs_eigenvalues(orc, t):
    kind, n, gold = LANCZS_GOLD[t]
    m = 7.49 * S.AMU
    x, sh = S.grid(n, 0.5, 2.8)
    shm = 1.0 / (24.0 * m * sh * sh)
    v0 = S.ohs(x, 1.146, 1943.79, m) if kind == "ohs" else S.morse_cm(x, 1.146, 1943.79, 13.71, m)
    nreort = min(n, 300)  # NREORT = LMTREORT (espec.f:151, param.h:9)
    info, ev, _ = orc.lanczs([n], 1, len(gold), nreort, [shm], v0, 1e-6)
    assert info == 0
    assert np.abs(ev - np.array(gold)).max() <= 6e-7
