"""The oracle's pulse-driven finite-difference leapfrog, PPSOD
(ppsod.f:1-548) with its step PSODAUX (sod.f:3-122), pinned:

- against the reference's own PPSOD compiled from its sources
  (oracle/_ref, skipped where that build is absent): C(t) samples, the
  final state and VPOT as the reference leaves it (the last VEFF), for
  every pulse shape, without absorber and with .VOPTIC and .SMOOTHW;
- by the mathematics: without a field PSODAUX is the leapfrog
  psi(t+dt) = psi(t-dt) - i cst H psi(t) with H the dense FD matrix
  (tests/golden/setups.py fd_matrix), the scheme conserves the leapfrog
  norm to within 1e-6 over a stable run, and a spatially constant dipole
  d0 multiplies C(t) by exp(i d0 cst sum E) to the scheme's accuracy,
  which fixes the sign of the coupling.
"""
import numpy as np
import pytest

from oracle import ABS_NONE, ABS_SMOOTHW, ABS_VOPTIC, SHFS

A0A, AMU = 0.5291772083, 1836.152667539


def case(n=201):
    xi, xf = 0.5 / A0A, 2.8 / A0A
    x = np.linspace(xi, xf, n)
    sh = x[1] - x[0]
    m = 7.49 * AMU
    xe = 1.146 / A0A
    vp = 0.5 * 0.35 * (x - xe) ** 2
    psi = np.exp(-((x - xe - 0.05) ** 2) / (2 * 0.06 ** 2))
    psi /= np.linalg.norm(psi)
    d = np.clip(np.maximum(0.3 - (x - xi) / (xf - xi), (x - xi) / (xf - xi) - 0.7), 0, None)
    return dict(np_=[n], shm=[1.0 / (24.0 * m * sh * sh)], vp=vp, u=psi, v=np.zeros(n),
                dm=0.8 * (x - xe), vopt=-0.3 * (d / 0.3) ** 2, mask=np.exp(-2.0 * d ** 2))


def pulse(efc):
    return (efc, 2e9, 0.05, 0.0, 0.06, 2.0, 0.3, 2)


@pytest.mark.parametrize("efc", [0, 1, 2, 3])
@pytest.mark.parametrize("absorb", [ABS_NONE, ABS_VOPTIC, ABS_SMOOTHW])
def test_ppsod_matches_reference(orc, ref, efc, absorb):
    c = case()
    va = {ABS_NONE: None, ABS_VOPTIC: c["vopt"], ABS_SMOOTHW: c["mask"]}[absorb]
    dt, tf, lc = 2e-5, 60 * 2e-5, 2
    o = orc.ppsod(c["np_"], dt, 0.0, tf, lc, c["shm"], c["u"], c["v"], c["vp"], c["dm"],
                  pulse(efc), absorb, va)
    r = ref.ppsod(c["np_"], dt, 0.0, tf, lc, c["shm"], c["u"], c["v"], c["vp"], c["dm"],
                  pulse(efc), absorb, va, nsamp=len(o["cr"]))
    assert r["info"] == 0
    assert max(np.abs(o["cr"] - r["cr"]).max(), np.abs(o["ci"] - r["ci"]).max()) <= 1e-13
    assert np.abs(o["u"] - r["u"]).max() <= 1e-13 and np.abs(o["v"] - r["v"]).max() <= 1e-13
    assert np.array_equal(o["veff"], r["veff"])


def test_psodaux_is_the_fd_leapfrog(orc):
    import setups as S
    c = case(64)
    h = S.fd_matrix(c["vp"], c["shm"][0])
    rng = np.random.default_rng(1)
    u1, u2, v1, v2 = (rng.standard_normal(64) for _ in range(4))
    cst = 0.3
    a1, a2, b1, b2 = orc.psodaux(c["np_"], cst, c["shm"], u1, u2, v1, v2, c["vp"])
    new = (u1 + 1j * v1) - 1j * cst * h @ (u2 + 1j * v2)
    assert np.abs(a2 + 1j * b2 - new).max() <= 1e-12
    assert np.array_equal(a1, u2) and np.array_equal(b1, v2)


def test_ppsod_without_field_is_unitary_enough_and_constant_dipole_phase(orc):
    c = case()
    dt, tf = 2e-5, 400 * 2e-5
    free = orc.ppsod(c["np_"], dt, 0.0, tf, 1, c["shm"], c["u"], c["v"], c["vp"], c["vp"] * 0,
                     pulse(0))
    assert np.abs(free["norm"] - 1.0).max() <= 1e-6
    d0 = 0.7
    f = orc.ppsod(c["np_"], dt, 0.0, tf, 1, c["shm"], c["u"], c["v"], c["vp"],
                  np.full(c["vp"].size, d0), pulse(1))
    # field at each step's end time (the start-up's first half-step at (T+DT)/2)
    cst = dt / SHFS
    t = np.concatenate([[dt], dt + dt * np.arange(1, len(f["cr"]) - 1)])
    e = np.array([orc.ef(*pulse(1)[:8], tt) for tt in t])
    phase = np.exp(1j * d0 * cst * np.concatenate([[0.0], np.cumsum(e)]))
    g = (free["cr"] + 1j * free["ci"]) * phase
    # the leapfrog takes each step's field at the end time of a two-step
    # update, so the phase agrees to O(dt) of the accumulated angle (0.1%
    # here); the opposite sign of the coupling would miss by twice the angle
    ang = np.abs(np.angle(phase)).max()
    assert ang > 1e-3
    assert np.abs((f["cr"] + 1j * f["ci"]) - g).max() <= 2e-3 * ang
