"""Pulse-driven Lanczos propagation (PPLNZ) on the GPU against the oracle.

Every step forms VEFF = VP - DM EF(T) on the device (CHPOT, chpot.f:58-62)
from the per-step field the host evaluates with EF (ef.f), then takes the
PLNZAUX step.  The oracle's EF and PPLNZ are pinned against the reference
build in tests/test_oracle_pulse.py.

Bars: |dC(t)|, |d psi| <= 1e-10 for FP64, <= 1e-6 for the mixed path.
"""
import numpy as np
import pytest

import espec_amd as E

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOLMIX = 1e-6


def case_1d(n=256):
    xi, xf = 0.5 / E.A0A, 2.8 / E.A0A
    x = np.linspace(xi, xf, n)
    sh = x[1] - x[0]
    m = 7.49 * E.AMU
    xe = 1.146 / E.A0A
    vp = 0.0125 * (x - xe) ** 2
    psi = np.exp(-((x - xe - 0.1) ** 2) / (2 * 0.05 ** 2))
    psi /= np.sqrt((psi ** 2).sum())
    return dict(np_=[n], shm=[1.0 / (24 * m * sh * sh)], vp=vp, u=psi, v=np.zeros(n),
                dm=0.8 * (x - xe))


def case_2d(np_=(64, 32), seed=2):
    rng = np.random.default_rng(seed)
    axes = [np.linspace(-2.0, 2.0, n) for n in np_]
    q1, q2 = np.meshgrid(axes[0], axes[1], indexing="xy")   # x fastest
    vp = 0.08 * q1 ** 2 + 0.05 * q2 ** 2 + 0.01 * q1 * q2
    psi = np.exp(-((q1 - rng.uniform(-0.3, 0.3)) ** 2 + (q2 - 0.2) ** 2) / 0.3)
    psi /= np.linalg.norm(psi)
    h = [a[1] - a[0] for a in axes]
    shm = [1.0 / (24 * 1836.0 * h[0] ** 2), 1.0 / (24 * 918.0 * h[1] ** 2)]
    return dict(np_=list(np_), shm=shm, vp=vp.ravel(), u=psi.ravel(), v=np.zeros(psi.size),
                dm=(0.6 * q1 - 0.3 * q2).ravel())


def pulse(efc):
    return (efc, 2e9, 0.05, 0.0, 0.06, 2.0, 0.3, 2)


def both(orc, c, mp, dt, tf, lc, pl, flags, mode):
    print(f"case: PPLNZ np={c['np_']} mp={mp} tf={tf} lc={lc} efc={pl[0]} flags={flags}",
          flush=True)
    g = E.pplnz(c["np_"], mp, 0.0, dt, 0.0, tf, lc, c["shm"], c["u"], c["v"], c["vp"], c["dm"],
                pl, flags)
    o = orc.pplnz(c["np_"], mp, 0.0, dt, 0.0, tf, lc, c["shm"], c["u"], c["v"], c["vp"],
                  c["dm"], pl, mode)
    assert len(g["cr"]) == len(o["cr"])
    assert np.array_equal(g["veff"], o["veff"])
    return g, o


def dC(a, b):
    return max(np.abs(a["cr"] - b["cr"]).max(), np.abs(a["ci"] - b["ci"]).max())


@pytest.mark.parametrize("efc", [E.PULSE_GAUS, E.PULSE_GGAUS, E.PULSE_SIN2])
@pytest.mark.parametrize("modename", ["REF", "EXACT"])
@pytest.mark.parametrize("lc", [1, 3])
def test_pplnz_1d_fp64(gpu, orc, efc, modename, lc):
    import oracle
    c = case_1d()
    g, o = both(orc, c, 10, 1e-3, 0.1, lc, pulse(efc), E.FP64 | getattr(E, modename),
                getattr(oracle, modename))
    assert dC(g, o) <= TOL64
    assert np.abs(g["u"] - o["u"]).max() <= TOL64 and np.abs(g["v"] - o["v"]).max() <= TOL64


@pytest.mark.parametrize("efc", [E.PULSE_GAUS, E.PULSE_SIN2])
@pytest.mark.parametrize("flags,tol", [(E.FP64 | E.EXACT, TOL64), (E.MIXED | E.EXACT, TOLMIX)])
def test_pplnz_2d(gpu, orc, efc, flags, tol):
    from oracle import EXACT
    c = case_2d()
    g, o = both(orc, c, 12, 2e-3, 0.12, 1, pulse(efc), flags, EXACT)
    assert dC(g, o) <= tol


def test_sil_plan_pulse_batch(gpu, orc):
    """Three packets with their own dipoles sharing one field in one plan."""
    from oracle import EXACT
    cs = [case_2d(seed=s) for s in (3, 4, 5)]
    dms = [cs[k]["dm"] * (1.0 + 0.5 * k) for k in range(3)]
    dt, nst, pl = 2e-3, 40, pulse(E.PULSE_GAUS)
    t, ef = 0.0, []
    for _ in range(nst):
        t += dt
        ef.append(E.ef(*pl, t))
    c0 = cs[0]
    P = E.SilPlan(c0["np_"], 12, 0.0, dt, 1, c0["shm"], np.stack([c["vp"] for c in cs]),
                  np.stack([c["u"] for c in cs]), None, maxsamp=nst + 1, flags=E.FP64 | E.EXACT)
    try:
        P.set_pulse(np.stack(dms), np.array(ef))
        P.run(nst)
        P.sync()
        with pytest.raises(E.EspecError):
            P.run(1)  # past the last field value
        for k, c in enumerate(cs):
            g = P.get(k)
            o = orc.pplnz(c0["np_"], 12, 0.0, dt, 0.0, nst * dt, 1, c0["shm"], c["u"], c["v"],
                          c["vp"], dms[k], pl, EXACT)
            n = len(o["cr"])
            assert dC({"cr": g["cr"][:n], "ci": g["ci"][:n]}, o) <= TOL64
    finally:
        P.close()


def test_pulse_plan_rejects_adaptive_steps(gpu):
    c = case_2d()
    P = E.SilPlan(c["np_"], 21, 0.0, 2e-3, 1, c["shm"], c["vp"], c["u"], None, maxsamp=8)
    try:
        P.set_pulse(c["dm"], np.zeros(4))
        with pytest.raises(E.EspecError) as ei:
            P.adaptive(2, None, tol=1e-8, m_fix=12)
        assert ei.value.code == 6
    finally:
        P.close()
