"""Split-operator propagation on the GPU against the FP64 oracle.

Bars: |dC(t)| <= 1e-10 for the FP64 path (ESPEC_FP64, reference or exact
semantics) and <= 1e-6 for the mixed path (complex64 state, FP64 C(t)).
"""
import numpy as np
import pytest

import espec_amd as E
from espec_amd import configs

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOLMIX = 1e-6


def harmonic_1d(n=256, lo=0.5, hi=2.8, mass_amu=7.49, x0_shift=0.05, sigma=0.06):
    """A test12-like bound packet (stable regime of the reference step)."""
    xi, xf = lo / E.A0A, hi / E.A0A
    x = np.linspace(xi, xf, n)
    h = x[1] - x[0]
    m = mass_amu * E.AMU
    xe = 1.146 / E.A0A
    vpot = 0.5 * 0.35 * (x - xe) ** 2
    psi = np.exp(-((x - xe - x0_shift) ** 2) / (2 * sigma ** 2))
    psi /= np.sqrt((psi ** 2).sum())
    return dict(np_=[n], sh=[h], sm=[m], vpot=vpot, u=psi, v=np.zeros(n))


def dC(a, b):
    return max(np.abs(a["cr"] - b["cr"]).max(), np.abs(a["ci"] - b["ci"]).max())


def run_both(orc, p, dt, ti, tf, lc, flags, mode):
    print(f"case: np={p['np_']} dt={dt} steps={E.nsteps(ti, tf, dt)} lc={lc} flags={flags}", flush=True)
    g = E.pspofft(p["np_"], dt, ti, tf, lc, p["sh"], p["sm"], p["u"], p["v"], p["vpot"], flags)
    o = orc.pspofft(p["np_"], dt, ti, tf, lc, p["sh"], p["sm"], p["u"], p["v"], p["vpot"], mode)
    assert len(g["cr"]) == len(o["cr"])
    return g, o


def test_ref_fp64_1d_matches_oracle(gpu, orc):
    from oracle import REF
    p = harmonic_1d()
    g, o = run_both(orc, p, 1e-4, 0.0, 0.2, 50, E.FP64 | E.REF, REF)
    assert dC(g, o) <= TOL64
    assert np.abs(g["u"] - o["u"]).max() <= TOL64
    assert np.abs(g["v"] - o["v"]).max() <= TOL64


@pytest.mark.parametrize("lc", [1, 3])
def test_ref_fp64_1d_sampling_rule(gpu, orc, lc):
    from oracle import REF
    p = harmonic_1d(n=128)
    g, o = run_both(orc, p, 2e-4, 0.0, 0.05, lc, E.FP64 | E.REF, REF)
    assert dC(g, o) <= TOL64


def test_exact_fp64_1d_config1(gpu, orc):
    from oracle import EXACT
    c = configs.config1()
    p = dict(np_=c.np, sh=c.sh, sm=c.sm, vpot=c.vpot, u=c.u, v=c.v)
    tf = 600 * c.dt_fs
    g, o = run_both(orc, p, c.dt_fs, 0.0, tf, 1, E.FP64 | E.EXACT, EXACT)
    assert dC(g, o) <= TOL64
    assert np.abs(g["norm"] - 1.0).max() < 1e-12


def test_mixed_1d_config1_full(gpu, orc):
    """Config 1 end to end (4096 steps, C(t) every step) in mixed precision."""
    from oracle import EXACT
    c = configs.config1()
    p = dict(np_=c.np, sh=c.sh, sm=c.sm, vpot=c.vpot, u=c.u, v=c.v)
    tf = (c.nsteps - 1) * c.dt_fs  # DO loop runs (TF-TI)/DT + 1 times
    g, o = run_both(orc, p, c.dt_fs, 0.0, tf, 1, E.MIXED | E.EXACT, EXACT)
    assert len(g["cr"]) == c.nsteps + 1
    err = dC(g, o)
    print("config1 mixed max|dC| =", err)
    assert err <= TOLMIX
    assert np.abs(g["norm"] - 1.0).max() < 1e-5


@pytest.mark.parametrize("flags,tol", [(E.FP64, TOL64), (E.MIXED, TOLMIX)])
def test_2d_lds_resident(gpu, orc, flags, tol):
    from oracle import EXACT
    c = configs.packet2d(64, [0.15, 0.12], [1.0, 1.2], [1200.0, 900.0], [0.3, 0.2], [0.2, 0.25],
                         1.0, 0)
    p = dict(np_=c.np, sh=c.sh, sm=c.sm, vpot=c.vpot, u=c.u, v=c.v)
    g, o = run_both(orc, p, c.dt_fs, 0.0, 150 * c.dt_fs, 2, flags | E.EXACT, EXACT)
    assert dC(g, o) <= tol


@pytest.mark.parametrize("flags,tol", [(E.FP64, TOL64), (E.MIXED, TOLMIX)])
def test_2d_multipass_config2(gpu, orc, flags, tol):
    """The 256x256 config-2 surface through the row/column pass kernels."""
    from oracle import EXACT
    c = configs.config2()
    p = dict(np_=c.np, sh=c.sh, sm=c.sm, vpot=c.vpot, u=c.u, v=c.v)
    g, o = run_both(orc, p, c.dt_fs, 0.0, 80 * c.dt_fs, 1, flags | E.EXACT, EXACT)
    assert dC(g, o) <= tol
    assert np.abs(g["norm"] - 1.0).max() < (1e-12 if flags == E.FP64 else 1e-5)


def test_3d_multipass(gpu, orc):
    from oracle import EXACT
    c = configs.config3(n=32)
    p = dict(np_=c.np, sh=c.sh, sm=c.sm, vpot=c.vpot, u=c.u, v=c.v)
    g, o = run_both(orc, p, c.dt_fs, 0.0, 20 * c.dt_fs, 1, E.FP64 | E.EXACT, EXACT)
    assert dC(g, o) <= TOL64


def test_batched_plan_matches_single_packets(gpu, orc):
    """Eight different 1-D packets in one plan; each against its own oracle run."""
    from oracle import EXACT
    rng = np.random.default_rng(7)
    pk = [configs.packet1d(256, rng.uniform(0.1, 0.2), rng.uniform(0.8, 1.4), 900.0,
                           rng.uniform(0.1, 0.4), 0.2, 0.5, 0) for _ in range(8)]
    vp = np.stack([p.vpot for p in pk])
    u = np.stack([p.u for p in pk])
    nst, lc = 300, 5
    for flags, tol in [(E.FP64, TOL64), (E.MIXED, TOLMIX)]:
        plan = E.SpoPlan([256], pk[0].dt_fs, lc, pk[0].sh, pk[0].sm, vp, u, None,
                         maxsamp=nst // lc + 1, flags=flags)
        plan.run(nst // 2)
        plan.run(nst - nst // 2)  # two runs continue the same propagation
        for k, p in enumerate(pk):
            g = plan.get(k)
            o = orc.pspofft(p.np, p.dt_fs, 0.0, (nst - 1) * p.dt_fs, lc, p.sh, p.sm, p.u, p.v,
                            p.vpot, EXACT)
            assert dC(g, o) <= tol, (flags, k)
        plan.close()
