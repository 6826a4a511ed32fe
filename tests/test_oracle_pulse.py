"""Pulse-driven Lanczos propagation in the oracle: EF (ef.f), CHPOT
(chpot.f:58-62) and PPLNZ's time loop (pplnz.f:72-300).

Pinned against the reference's own EF and PPLNZ compiled from its sources
(oracle/_ref; skipped where absent), and by mathematics: without a field
PPLNZ is PLNZ, and a spatially constant dipole d0 only shifts H by
-d0 E(T), so every C(t) sample gains the phase exp(i d0 cst sum_g E_g),
which fixes the sign of the coupling and the time at which the field is
taken (the step's end time).
"""
import numpy as np
import pytest

from oracle import (EXACT, PULSE_GAUS, PULSE_GGAUS, PULSE_NONE, PULSE_SIN2, REF, SHFS)

A0A = 0.5291772083
AMU = 1836.152667539
KINDS = [PULSE_NONE, PULSE_GAUS, PULSE_GGAUS, PULSE_SIN2]


@pytest.fixture(scope="module")
def case():
    n = 256
    xi, xf = 0.5 / A0A, 2.8 / A0A
    sh = (xf - xi) / (n - 1)
    x = xi + sh * np.arange(n)
    m = 7.49 * AMU
    xe = 1.146 / A0A
    vpot = 0.0125 * (x - xe) ** 2
    psi = np.exp(-((x - xe - 0.1) ** 2) / (2 * 0.05 ** 2))
    psi /= np.sqrt((psi ** 2).sum())
    return dict(n=n, x=x, vp=vpot, u=psi, v=np.zeros(n), dm=0.8 * (x - xe),
                shm=[1.0 / (24 * m * sh * sh)])


def pulse(efc, kl=2):
    # E0, T0, TD, TP (fs), OMG (eV), SNI, KL: a few-fs pulse around 0.05 fs
    return (efc, 2e9, 0.05, 0.0, 0.06, 2.0, 0.3, kl)


@pytest.mark.parametrize("efc", KINDS)
@pytest.mark.parametrize("kl", [1, 2, 3])
def test_ef_matches_reference(orc, ref, efc, kl):
    for t in (0.0, 0.02, 0.05, 0.09, 0.5, 3.0, 12.0):
        args = (efc, 2e9, 0.05, 0.01, 0.06, 2.0, 0.3, kl, t)
        assert orc.ef(*args) == ref.ef(*args)


@pytest.mark.parametrize("efc", KINDS)
@pytest.mark.parametrize("lc", [1, 3])
def test_pplnz_matches_reference(orc, ref, case, efc, lc):
    c = case
    args = ([c["n"]], 10, 0.0, 1e-3, 0.0, 0.1, lc, c["shm"], c["u"], c["v"], c["vp"], c["dm"],
            pulse(efc))
    r = ref.pplnz(*args)
    o = orc.pplnz(*args, REF)
    k = len(o["cr"])
    assert len(r["cr"]) == k == orc.nsteps(0.0, 0.1 - 1e-3 / 2, 1e-3) // lc + 1
    # With PRTCRL='.NO' the reference never stores sample 0 (pplnz.f:130-134).
    assert r["cr"][0] == 0.0 and r["ci"][0] == 0.0
    assert np.abs(r["cr"][1:] - o["cr"][1:]).max() <= 1e-13
    assert np.abs(r["ci"][1:] - o["ci"][1:]).max() <= 1e-13
    assert np.abs(r["u"] - o["u"]).max() <= 1e-13 and np.abs(r["v"] - o["v"]).max() <= 1e-13
    assert np.array_equal(r["veff"], o["veff"])
    if efc != PULSE_NONE:  # the field acted
        p = orc.plnz([c["n"]], 10, 0.0, 1e-3, 0.0, 0.1 - 1e-3, lc, c["shm"], c["u"], c["v"],
                     c["vp"], REF)
        assert np.abs(p["cr"][:k] - o["cr"]).max() > 1e-7


@pytest.mark.parametrize("mode", [REF, EXACT])
def test_no_field_is_plnz(orc, case, mode):
    c = case
    dt, k = 1e-3, 60
    o = orc.pplnz([c["n"]], 10, 0.0, dt, 0.0, k * dt, 1, c["shm"], c["u"], c["v"], c["vp"],
                  c["dm"], pulse(PULSE_NONE), mode)
    p = orc.plnz([c["n"]], 10, 0.0, dt, 0.0, (k - 1) * dt, 1, c["shm"], c["u"], c["v"], c["vp"],
                 mode)
    assert len(o["cr"]) == len(p["cr"]) == k + 1
    assert np.array_equal(o["cr"], p["cr"]) and np.array_equal(o["u"], p["u"])
    assert np.array_equal(o["veff"], c["vp"])


@pytest.mark.parametrize("efc", [PULSE_GAUS, PULSE_SIN2])
def test_constant_dipole_is_a_phase(orc, case, efc):
    """d0 constant: V_eff = V - d0 E(T_g) in step g, so
    C_k = C_k(no field) exp(i d0 cst sum_{g<=k} E(T_g)) with T_g = TI + g dt."""
    c = case
    dt, k, d0 = 1e-3, 80, 0.7
    pl = pulse(efc)
    o = orc.pplnz([c["n"]], 12, 0.0, dt, 0.0, k * dt, 1, c["shm"], c["u"], c["v"], c["vp"],
                  np.full(c["n"], d0), pl, EXACT)
    p = orc.plnz([c["n"]], 12, 0.0, dt, 0.0, (k - 1) * dt, 1, c["shm"], c["u"], c["v"], c["vp"],
                 EXACT)
    t, es = 0.0, [0.0]
    for g in range(k):
        t += dt
        es.append(orc.ef(*pl[:7], pl[7], t))
    phase = np.exp(1j * d0 * dt / SHFS * np.cumsum(es))
    got = o["cr"] + 1j * o["ci"]
    want = (p["cr"] + 1j * p["ci"]) * phase
    assert np.abs(np.angle(phase[-1])) > 1e-3  # the field moved the phase
    assert np.abs(got - want).max() <= 1e-12
