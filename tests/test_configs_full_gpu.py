"""The benchmark surfaces at their full sizes and lengths, on the GPU,
against the FP64 oracle (oracle/espec_oracle.c).

* 1-D Morse, 512 points, 4096 steps of dt = 0.5 a.u., FP64 path.
* 256x256 coupled Morse, 2000 steps of dt = 1 a.u., mixed path.
* 256x256 short-iterative Lanczos, fixed orders 10 and 12, ABSTOL <= 0
  (QL eigensolve, dsteqr.f semantics), FP64 and mixed, 50 steps.
* 128^3 three-mode surface: 20 steps of both paths against the oracle,
  then all 1000 steps of the mixed path against the FP64 path.
* The 2048-packet uneven batch, 1000 steps: one packet of every grid class
  against its own oracle run over all 1000 steps.

Bars (|C(t) - C_oracle(t)| over every sample): 1e-10 for the FP64 path,
1e-6 for the mixed path (complex64 state, FP64 norm and C(t)).  The
oracle follows pspofft.f:110-146 / spofft.f:3-203 (EXACT mode) and
plnz.f:109-150 / sil.f:70-181.
"""
import numpy as np
import pytest

import espec_amd as E
from espec_amd import configs, krylov

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOLMIX = 1e-6


def dC(a, b, n=None):
    n = n or min(len(a["cr"]), len(b["cr"]))
    return max(np.abs(a["cr"][:n] - b["cr"][:n]).max(), np.abs(a["ci"][:n] - b["ci"][:n]).max())


def spo(c, flags, nsteps):
    return E.pspofft(c.np, c.dt_fs, 0.0, (nsteps - 1) * c.dt_fs, 1, c.sh, c.sm, c.u, c.v,
                     c.vpot, flags)


def spo_oracle(orc, c, nsteps):
    from oracle import EXACT
    return orc.pspofft(c.np, c.dt_fs, 0.0, (nsteps - 1) * c.dt_fs, 1, c.sh, c.sm, c.u, c.v,
                       c.vpot, EXACT)


def test_config1_fp64_all_4096_steps(gpu, orc):
    c = configs.config1()
    print(f"case: config1 FP64 {c.nsteps} steps", flush=True)
    g, o = spo(c, E.FP64 | E.EXACT, c.nsteps), spo_oracle(orc, c, c.nsteps)
    assert len(g["cr"]) == len(o["cr"]) == c.nsteps + 1
    err = dC(g, o)
    print(f"config1 FP64 max|dC| {err:.2e}", flush=True)
    assert err <= TOL64
    assert np.abs(g["norm"] - 1.0).max() < 1e-12


def test_config2_spo_mixed_all_2000_steps(gpu, orc):
    c = configs.config2()
    print(f"case: config2 SO-FFT mixed {c.nsteps} steps", flush=True)
    g, o = spo(c, E.MIXED | E.EXACT, c.nsteps), spo_oracle(orc, c, c.nsteps)
    assert len(g["cr"]) == c.nsteps + 1
    err = dC(g, o)
    drift = np.abs(g["norm"] - 1.0).max()
    print(f"config2 mixed max|dC| {err:.2e} norm drift {drift:.1e}", flush=True)
    assert err <= TOLMIX
    assert drift < 1e-5


def plnz_pair(orc, c, mp, flags, u=None):
    from oracle import EXACT
    shm = krylov.shm_of(c)
    tf = (c.nsteps - 1) * c.dt_fs
    u = c.u if u is None else u
    print(f"case: config2 PLNZ order {mp} {c.nsteps} steps flags={flags}", flush=True)
    g = E.plnz(c.np, mp, 0.0, c.dt_fs, 0.0, tf, 1, shm, u, c.v, c.vpot, flags)
    o = orc.plnz(c.np, mp, 0.0, c.dt_fs, 0.0, tf, 1, shm, u, c.v, c.vpot, EXACT)
    assert len(g["cr"]) == len(o["cr"]) == c.nsteps + 1
    return g, o


@pytest.mark.parametrize("mp", [10, 12])
def test_config2_lanczos_fp64(gpu, orc, mp):
    """The Lanczos half of the 256x256 workload in FP64, ABSTOL = 0 (QL
    eigensolve), dt = 1 a.u.

    At order 10 this propagation amplifies rounding: the oracle itself,
    started from psi0 perturbed by 1e-16 (relative, random), departs from
    its own unperturbed run by about 1e-11 after 20 steps and 3e-9 after
    50 (order 12: 3e-15 and 8e-13).  So order 10 is held to 1e-10 over 20
    steps, and over 50 steps to four times that measured sensitivity,
    which the test measures again on the same inputs; order 12 is held to
    1e-10 over all 50 steps."""
    c20 = configs.config2(nsteps=20)
    if mp == 10:
        g, o = plnz_pair(orc, c20, mp, E.FP64 | E.EXACT)
        err = dC(g, o)
        print(f"order {mp}, 20 steps: max|dC| {err:.2e}", flush=True)
        assert err <= TOL64
    c = configs.config2(nsteps=50)
    g, o = plnz_pair(orc, c, mp, E.FP64 | E.EXACT)
    err = dC(g, o)
    bar = TOL64
    if mp == 10:
        from oracle import EXACT
        rng = np.random.default_rng(1)
        up = c.u * (1.0 + 1e-16 * rng.standard_normal(c.u.size))
        op = orc.plnz(c.np, mp, 0.0, c.dt_fs, 0.0, (c.nsteps - 1) * c.dt_fs, 1,
                      krylov.shm_of(c), up, c.v, c.vpot, EXACT)
        sens = dC(o, op)
        bar = max(TOL64, 4.0 * sens)
        print(f"oracle sensitivity to a 1e-16 start perturbation: {sens:.2e}", flush=True)
    print(f"order {mp}, 50 steps: max|dC| {err:.2e} (bar {bar:.1e})", flush=True)
    assert err <= bar
    if mp == 12:
        assert np.abs(g["u"] - o["u"]).max() <= TOL64 and np.abs(g["v"] - o["v"]).max() <= TOL64


def test_config2_lanczos_order10_mixed(gpu, orc):
    g, o = plnz_pair(orc, configs.config2(nsteps=50), 10, E.MIXED | E.EXACT)
    err = dC(g, o)
    print(f"order 10 mixed, 50 steps: max|dC| {err:.2e}", flush=True)
    assert err <= TOLMIX


@pytest.fixture(scope="module")
def config3_oracle_20(orc):
    c = configs.config3(nsteps=20)
    return c, spo_oracle(orc, c, c.nsteps)


@pytest.mark.parametrize("flags,tol", [(E.FP64 | E.EXACT, TOL64), (E.MIXED | E.EXACT, TOLMIX)])
def test_config3_128cubed_vs_oracle(gpu, config3_oracle_20, flags, tol):
    c, o = config3_oracle_20
    print(f"case: config3 128^3 {c.nsteps} steps flags={flags}", flush=True)
    g = spo(c, flags, c.nsteps)
    assert len(g["cr"]) == c.nsteps + 1
    err = dC(g, o)
    print(f"config3 max|dC| {err:.2e}", flush=True)
    assert err <= tol


def test_config3_mixed_vs_fp64_all_1000_steps(gpu):
    c = configs.config3()
    print(f"case: config3 128^3 mixed vs FP64 {c.nsteps} steps", flush=True)
    g64 = spo(c, E.FP64 | E.EXACT, c.nsteps)
    gmx = spo(c, E.MIXED | E.EXACT, c.nsteps)
    err = dC(gmx, g64)
    print(f"config3 mixed vs FP64 max|dC| {err:.2e}", flush=True)
    assert err <= TOLMIX
    assert np.abs(g64["norm"] - 1.0).max() < 1e-11
    assert np.abs(gmx["norm"] - 1.0).max() < 1e-5


def test_config4_batch_all_1000_steps(gpu, orc):
    """All 2048 packets for 1000 steps in the batched plans; one packet of
    every grid class against its own FP64 oracle run over the 1000 steps."""
    from espec_amd.batch import SpoBatch
    steps = 1000
    pk = configs.config4(count=2048, nsteps=steps, seed=4)
    picks = {}
    for i, p in enumerate(pk):
        picks.setdefault(tuple(p.np), i)
    print(f"case: config4 batch 2048 packets {steps} steps, checked {sorted(picks)}", flush=True)
    b = SpoBatch(pk, flags=E.MIXED | E.EXACT)
    b.run(steps)
    b.sync()
    for key, i in sorted(picks.items()):
        p = pk[i]
        o = spo_oracle(orc, p, steps)
        g = b.get(i)
        assert len(g["cr"]) == steps + 1
        err = dC(g, o)
        print(f"config4 packet {i} grid {key} max|dC| {err:.2e}", flush=True)
        assert err <= TOLMIX, (key, err)
    b.close()
