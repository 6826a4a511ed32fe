"""Short-iterative Lanczos propagation (PLNZ) on the GPU against the oracle.

Bars: |dC(t)|, |d psi| <= 1e-10 for the FP64 path with ABSTOL <= 0 (QL
eigensolve, continuous in its inputs); <= 1e-6 for the mixed path
(complex64 state and Krylov basis, FP64 arithmetic).  With ABSTOL > 0 the
eigenvalues are midpoints of bisection intervals (dstebz.f), a piecewise
constant function of the tridiagonal, so rounding-level differences can move
an eigenvalue by up to ABSTOL; that case is held to ABSTOL * cst per step.
"""
import os
import sys

import numpy as np
import pytest

import espec_amd as E

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import setups  # noqa: E402

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOLMIX = 1e-6


def ohs_case(n=256):
    """The OHS pair of the reference's tests 12/14 (201/256-point grid on
    [0.5, 2.8] A, 7.49 amu): ground state of the lower well on the upper."""
    c = setups.test_case(12)
    if n != 256:
        m = c["sm"][0]
        x, sh = setups.grid(n, 0.5, 2.8)
        shm = 1.0 / (24.0 * m * sh * sh)
        v0 = setups.ohs(x, 1.146, 1943.79, m)
        c.update(np_=[n], sh=[sh], shm=[shm], vpot=setups.ohs(x, 1.211, 1629.23, m),
                 u=setups.ground_state(v0, shm), v=np.zeros(n))
    return c


def coupled_2d(n0=32, n1=16, seed=3):
    """Two coupled anharmonic modes on a 32 x 16 grid, displaced Gaussian."""
    q0 = np.linspace(-3.0, 3.0, n0)
    q1 = np.linspace(-2.5, 2.5, n1)
    h = [q0[1] - q0[0], q1[1] - q1[0]]
    m = [1836.0, 918.0]
    Q0, Q1 = np.meshgrid(q0, q1, indexing="xy")  # x (axis 0) fastest
    vpot = 0.02 * Q0 ** 2 + 0.01 * Q1 ** 2 + 0.002 * Q0 ** 3 + 0.003 * Q0 * Q1
    psi = np.exp(-((Q0 - 0.4) ** 2) / 0.5 - ((Q1 + 0.2) ** 2) / 0.6)
    psi /= np.sqrt((psi ** 2).sum())
    shm = [1.0 / (24.0 * m[a] * h[a] ** 2) for a in range(2)]
    return dict(np_=[n0, n1], shm=shm, vpot=vpot.ravel(), u=psi.ravel(), v=np.zeros(n0 * n1))


def dC(a, b):
    return max(np.abs(a["cr"] - b["cr"]).max(), np.abs(a["ci"] - b["ci"]).max())


def run_both(orc, c, mp, abstol, dt, nst, lc, flags, mode):
    tf = (nst - 1) * dt
    print(f"case: PLNZ np={c['np_']} mp={mp} abstol={abstol} dt={dt} steps={nst} lc={lc} "
          f"flags={flags}", flush=True)
    g = E.plnz(c["np_"], mp, abstol, dt, 0.0, tf, lc, c["shm"], c["u"], c["v"], c["vpot"], flags)
    o = orc.plnz(c["np_"], mp, abstol, dt, 0.0, tf, lc, c["shm"], c["u"], c["v"], c["vpot"], mode)
    assert len(g["cr"]) == len(o["cr"]) == nst // lc + 1
    return g, o


@pytest.mark.parametrize("modename", ["REF", "EXACT"])
def test_plnz_fp64_ql_1d(gpu, orc, modename):
    import oracle
    mode = getattr(oracle, modename)
    flags = E.FP64 | (E.REF if modename == "REF" else E.EXACT)
    g, o = run_both(orc, ohs_case(), 10, 0.0, 2e-3, 200, 1, flags, mode)
    assert dC(g, o) <= TOL64
    assert np.abs(g["u"] - o["u"]).max() <= TOL64
    assert np.abs(g["v"] - o["v"]).max() <= TOL64
    # |psi|^2 from the device C(t) sampler agrees with the final state
    assert abs(g["norm"][-1] - np.sum(g["u"] ** 2 + g["v"] ** 2)) <= 1e-12


@pytest.mark.parametrize("lc", [2, 5])
def test_plnz_sampling_rule(gpu, orc, lc):
    from oracle import REF
    g, o = run_both(orc, ohs_case(128), 8, 0.0, 2e-3, 40, lc, E.FP64 | E.REF, REF)
    assert dC(g, o) <= TOL64


def test_plnz_mixed_1d(gpu, orc):
    from oracle import EXACT
    g, o = run_both(orc, ohs_case(), 10, 0.0, 2e-3, 400, 1, E.MIXED | E.EXACT, EXACT)
    assert dC(g, o) <= TOLMIX


def test_plnz_bisection_1d(gpu, orc):
    from oracle import REF
    dt, nst, abstol = 2e-3, 100, 1e-6
    g, o = run_both(orc, ohs_case(), 10, abstol, dt, nst, 1, E.FP64 | E.REF, REF)
    cst = dt / E.SHFS
    assert dC(g, o) <= abstol * cst * nst


@pytest.mark.parametrize("flags,tol", [(E.FP64 | E.EXACT, TOL64), (E.MIXED | E.EXACT, TOLMIX)])
def test_plnz_2d(gpu, orc, flags, tol):
    from oracle import EXACT
    g, o = run_both(orc, coupled_2d(), 12, 0.0, 0.05, 100, 1, flags, EXACT)
    assert dC(g, o) <= tol


def test_sil_plan_batched(gpu, orc):
    """Four packets with different potentials and starts in one plan, run
    in two pieces; each matches its own oracle propagation."""
    from oracle import REF
    base = ohs_case(128)
    n = 128
    rng = np.random.default_rng(7)
    vp = np.stack([base["vpot"] * (1 + 0.05 * k) for k in range(4)])
    u = np.stack([np.roll(base["u"], k) for k in range(4)])
    v = 0.01 * rng.standard_normal((4, n))
    dt, nst, mp = 2e-3, 30, 8
    print(f"case: sil plan npk=4 n={n} mp={mp} steps={nst}", flush=True)
    P = E.SilPlan([n], mp, 0.0, dt, 1, base["shm"], vp, u, v, maxsamp=nst + 1,
                  flags=E.FP64 | E.REF)
    P.run(12)
    P.run(nst - 12)
    for k in range(4):
        g = P.get(k)
        o = orc.plnz([n], mp, 0.0, dt, 0.0, (nst - 1) * dt, 1, base["shm"], u[k], v[k], vp[k], REF)
        assert dC(g, o) <= TOL64, k
        assert np.abs(g["u"] - o["u"]).max() <= TOL64
    P.close()


@pytest.mark.parametrize("flags", [E.FP64 | E.EXACT, E.MIXED | E.EXACT])
def test_no_graph_runs_match_graph_runs(flags):
    """ESPEC_NO_GRAPH launches the same kernels in the same order as the
    step graphs: identical C(t) and state, bit for bit, over a run that
    the graphs cover with one 8-step and several 1-step replays."""
    c = coupled_2d()
    dt = 0.05
    outs = []
    for extra in (0, E.NO_GRAPH):
        p = E.SilPlan(c["np_"], 8, 0.0, dt, 1, c["shm"], c["vpot"], c["u"], c["v"],
                      maxsamp=32, flags=flags | extra)
        p.run(11)
        p.run(3)
        p.sync()
        outs.append(p.get(0))
        p.close()
    a, b = outs
    assert len(a["cr"]) == len(b["cr"]) == 15
    for k in ("cr", "ci", "norm", "u", "v"):
        assert np.array_equal(a[k], b[k]), k
