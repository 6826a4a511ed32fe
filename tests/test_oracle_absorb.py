"""Absorbing boundaries in the oracle (ABSORB=.TRUE. with TPABSOR .VOPTIC or
.SMOOTHW): PSPOFFTAUX (spofft.f:65-118, 162-176) and PSODFFTAUX
(sodfft.f:66-72, 108-116).

Pinned three ways:
- against the reference's own PSPOFFT/PSODFFT compiled from its sources
  (oracle/_ref, skipped where that build is absent), in ORC_REF mode;
- by exact reductions: a zero imaginary potential or a unit mask is the
  step without absorber, bit for bit; a constant imaginary potential W
  scales every C(t) sample k by exp(W dt k) and a constant mask m by m^k
  (both commute with the whole step);
- by index: a mask that differs from 1 at a single point changes exactly
  that point.
The SOD mask branch (sodfft.f:108) compares TPABSOR(1:7) with the
8-character '.SMOOTHW' and is never taken, so the reference's PSODFFT
with .SMOOTHW propagates without absorber; ORC_REF keeps that, ORC_EXACT
applies each point's mask.  No published regression test of the reference
uses an absorber.
"""
import numpy as np
import pytest

from oracle import ABS_NONE, ABS_SMOOTHW, ABS_VOPTIC, EXACT, REF, SHFS

A0A = 0.5291772083
AMU = 1836.152667539


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
    # Imaginary potential rising quadratically over the outer 30% of each
    # side, and a smooth-walls mask of the same shape (absorb.f:63-75).
    d = np.clip(np.maximum(0.3 - (x - xi) / (xf - xi), (x - xi) / (xf - xi) - 0.7), 0, None)
    w = -0.8 * (d / 0.3) ** 2
    mask = np.exp(-3.0 * d ** 2)
    # a second, asymmetric pair that touches the packet's core
    w2 = -0.02 * (1.0 + np.sin(3.0 * x))
    mask2 = 1.0 - 2e-3 * (1.0 + np.cos(2.0 * x))
    return dict(n=n, sh=sh, m=m, vpot=vpot, u=psi, v=np.zeros(n), vopt=(w, w2),
                mask=(mask, mask2))


def cplx_err(a, b, c, d):
    return np.abs((a + 1j * b) - (c + 1j * d)).max()


def _vabc(c, absorb, k):
    return (c["vopt"] if absorb == ABS_VOPTIC else c["mask"])[k]


@pytest.mark.parametrize("absorb", [ABS_VOPTIC, ABS_SMOOTHW])
@pytest.mark.parametrize("k", [0, 1])
@pytest.mark.parametrize("lc", [1, 7])
def test_pspofft_abs_vs_reference(orc, ref, case, absorb, k, lc):
    c = case
    va = _vabc(c, absorb, k)
    args = ([c["n"]], 1e-3, 0.0, 0.5, lc, [c["sh"]], [c["m"]], c["u"], c["v"], c["vpot"])
    r = ref.pspofft_abs(*args, absorb, va)
    o = orc.pspofft_abs(*args, absorb, va, REF)
    kk = len(o["cr"])
    assert kk == orc.nsteps(0.0, 0.5, 1e-3) // lc + 1
    assert cplx_err(r["cr"][:kk], r["ci"][:kk], o["cr"], o["ci"]) <= 1e-12
    assert cplx_err(r["u"], r["v"], o["u"], o["v"]) <= 1e-12
    # the absorber did act: the norm fell below the unabsorbed run's
    plain = orc.pspofft(*args, REF)
    assert np.sum(o["u"] ** 2 + o["v"] ** 2) < np.sum(plain["u"] ** 2 + plain["v"] ** 2) - 1e-6


@pytest.mark.parametrize("absorb", [ABS_VOPTIC, ABS_SMOOTHW])
@pytest.mark.parametrize("k", [0, 1])
@pytest.mark.parametrize("lc", [1, 7])
def test_psodfft_abs_vs_reference(orc, ref, case, absorb, k, lc):
    c = case
    va = _vabc(c, absorb, k)
    args = ([c["n"]], 1e-3, 0.0, 0.5, lc, [c["sh"]], [c["m"]], c["u"], c["v"], c["vpot"])
    r = ref.psodfft_abs(*args, absorb, va)
    o = orc.psodfft_abs(*args, absorb, va, REF)
    kk = len(o["cr"])
    assert cplx_err(r["cr"][:kk], r["ci"][:kk], o["cr"], o["ci"]) <= 1e-13
    assert np.all(r["cr"][kk:] == 0)


@pytest.mark.parametrize("mode", [REF, EXACT])
def test_neutral_absorbers_are_the_plain_step(orc, case, mode):
    """W = 0 and a unit mask reproduce the runs without absorber exactly."""
    c = case
    args = ([c["n"]], 1e-3, 0.0, 0.2, 3, [c["sh"]], [c["m"]], c["u"], c["v"], c["vpot"])
    p = orc.pspofft(*args, mode)
    s = orc.psodfft(*args, mode)
    for absorb, va in ((ABS_VOPTIC, np.zeros(c["n"])), (ABS_SMOOTHW, np.ones(c["n"]))):
        a = orc.pspofft_abs(*args, absorb, va, mode)
        assert np.array_equal(a["cr"], p["cr"]) and np.array_equal(a["ci"], p["ci"])
        assert np.array_equal(a["u"], p["u"]) and np.array_equal(a["v"], p["v"])
        b = orc.psodfft_abs(*args, absorb, va, mode)
        assert np.array_equal(b["cr"], s["cr"]) and np.array_equal(b["u"], s["u"])


def test_spo_constant_absorbers_scale_c(orc, case):
    """A constant W multiplies psi by exp(cst2 W) in each half phase, i.e.
    C_k by exp(W dt k); a constant mask m multiplies C_k by m^k."""
    c = case
    dt = 2e-3
    args = ([c["n"]], dt, 0.0, 40 * dt, 1, [c["sh"]], [c["m"]], c["u"], c["v"], c["vpot"])
    p = orc.pspofft(*args, EXACT)
    k = np.arange(len(p["cr"]))
    w = -0.05
    a = orc.pspofft_abs(*args, ABS_VOPTIC, np.full(c["n"], w), EXACT)
    f = np.exp(w * dt / SHFS * k)
    assert np.abs(a["cr"] + 1j * a["ci"] - f * (p["cr"] + 1j * p["ci"])).max() <= 1e-13
    m = 0.995
    b = orc.pspofft_abs(*args, ABS_SMOOTHW, np.full(c["n"], m), EXACT)
    assert np.abs(b["cr"] + 1j * b["ci"] - m ** k * (p["cr"] + 1j * p["ci"])).max() <= 1e-13
    # and a sign slip in W would grow the norm instead
    assert np.abs(a["cr"][-1]) < np.abs(p["cr"][-1] + 1j * p["ci"][-1])


@pytest.mark.parametrize("where", ["first_half", "second_half"])
def test_sod_mask_index(orc, case, where):
    """ORC_EXACT: a mask differing from 1 only at point j changes exactly j;
    ORC_REF: nothing changes (the reference never reaches its mask loop)."""
    c = case
    n, nh = c["n"], c["n"] // 2
    tc, b = orc.digt([n], [c["sh"]], [c["m"]], EXACT)
    rng = np.random.default_rng(7)
    u1, u2, v1, v2 = (rng.standard_normal(n) for _ in range(4))
    j = 37 if where == "first_half" else nh + 37
    mask = np.ones(n)
    mask[j] = 0.5
    cst = 2 * 1e-3 / SHFS
    for mode in (REF, EXACT):
        base = orc.sod_step_abs([n], b, cst, u1, u2, v1, v2, tc, c["vpot"], ABS_NONE,
                                np.zeros(n), mode)
        got = orc.sod_step_abs([n], b, cst, u1, u2, v1, v2, tc, c["vpot"], ABS_SMOOTHW, mask,
                               mode)
        assert np.array_equal(got[0], base[0]) and np.array_equal(got[2], base[2])
        changed = set(np.nonzero((got[1] != base[1]) | (got[3] != base[3]))[0].tolist())
        want = {j} if mode == EXACT else set()
        assert changed == want
        for i in want:
            assert got[1][i] == 0.5 * base[1][i] and got[3][i] == 0.5 * base[3][i]


def test_sod_smoothw_reference_is_unabsorbed(orc, ref, case):
    """The reference's PSODFFT with .SMOOTHW equals its run without absorber."""
    c = case
    args = ([c["n"]], 1e-3, 0.0, 0.2, 1, [c["sh"]], [c["m"]], c["u"], c["v"], c["vpot"])
    r = ref.psodfft_abs(*args, ABS_SMOOTHW, c["mask"][1])
    p = ref.psodfft(*args)
    assert np.array_equal(r["cr"], p["cr"]) and np.array_equal(r["u"], p["u"])


def test_spo_mask_index(orc, case):
    """The SPO mask (spofft.f:171-174) scales exactly the masked point."""
    c = case
    n = c["n"]
    tc, b = orc.digt([n], [c["sh"]], [c["m"]], EXACT)
    rng = np.random.default_rng(8)
    u, v = rng.standard_normal(n), rng.standard_normal(n)
    cst1, cst2 = 1e-3 / SHFS, 0.5e-3 / SHFS
    for j in (5, n // 2 + 5):
        mask = np.ones(n)
        mask[j] = 0.25
        for mode in (REF, EXACT):
            bu, bv = orc.spo_step_abs([n], b, cst1, cst2, u, v, tc, c["vpot"], ABS_NONE,
                                      np.zeros(n), mode)
            gu, gv = orc.spo_step_abs([n], b, cst1, cst2, u, v, tc, c["vpot"], ABS_SMOOTHW,
                                      mask, mode)
            changed = np.nonzero((gu != bu) | (gv != bv))[0].tolist()
            assert changed == [j]
            assert gu[j] == 0.25 * bu[j] and gv[j] == 0.25 * bv[j]


# ------------------------------------------------------------ PLNZ/PLNZAUX
# PLNZAUX multiplies the new state by VABC when ABSORB and TPABSOR is
# .SMOOTHW (sil.f:171-180); PLNZ samples C(t) from that masked state.

def _plnz_case(c):
    shm = [1.0 / (24.0 * c["m"] * c["sh"] ** 2)]
    return shm, 2e-3, 60


@pytest.mark.parametrize("k", [0, 1])
def test_plnz_smoothw_matches_reference(orc, ref, case, k):
    c = case
    shm, dt, nst = _plnz_case(c)
    mask = c["mask"][k]
    tf = (nst - 1) * dt
    o = orc.plnz([c["n"]], 10, 0.0, dt, 0.0, tf, 1, shm, c["u"], c["v"], c["vpot"], REF,
                 mask=mask)
    r = ref.plnz_smoothw([c["n"]], 10, 0.0, dt, 0.0, tf, 1, shm, c["u"], c["v"], c["vpot"],
                         mask)
    assert r["info"] == 0
    assert cplx_err(o["cr"], o["ci"], r["cr"], r["ci"]) <= 1e-12
    assert np.abs(o["u"] - r["u"]).max() <= 1e-12 and np.abs(o["v"] - r["v"]).max() <= 1e-12
    # the mask touches the run: it differs from the run without absorber
    p = orc.plnz([c["n"]], 10, 0.0, dt, 0.0, tf, 1, shm, c["u"], c["v"], c["vpot"], REF)
    assert cplx_err(o["cr"], o["ci"], p["cr"], p["ci"]) > 1e-9


def test_plnz_smoothw_reductions(orc, case):
    """A unit mask is the plain step bit for bit; a constant mask m scales
    sample k by m^k (the step is linear in the state)."""
    c = case
    shm, dt, nst = _plnz_case(c)
    tf = (nst - 1) * dt
    args = ([c["n"]], 10, 0.0, dt, 0.0, tf, 1, shm, c["u"], c["v"], c["vpot"], EXACT)
    p = orc.plnz(*args)
    one = orc.plnz(*args, mask=np.ones(c["n"]))
    assert np.array_equal(p["cr"], one["cr"]) and np.array_equal(p["u"], one["u"])
    m = 0.999
    q = orc.plnz(*args, mask=np.full(c["n"], m))
    k = np.arange(len(p["cr"]))
    assert np.abs((q["cr"] + 1j * q["ci"]) - m ** k * (p["cr"] + 1j * p["ci"])).max() <= 1e-13
