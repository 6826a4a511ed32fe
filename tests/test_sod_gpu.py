"""Second-order-differencing propagation (PSODFFT) on the GPU against the oracle.

Bars: |dC(t)|, |d psi| <= 1e-10 for FP64 (reference or exact semantics),
<= 1e-6 for the mixed path.  Grids up to 2048 points run LDS-resident;
larger ones run the per-axis multi-pass path (exact semantics only).
"""
import numpy as np
import pytest

import espec_amd as E
import setups as S

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOLMIX = 1e-6


def morse_case(n=256):
    """The Morse pair of the reference's tests 5/6 on a 256-point grid."""
    c = S.test_case(5)
    if n != 256:
        m = c["sm"][0]
        x, sh = S.grid(n, 0.5, 2.8)
        shm = 1.0 / (24.0 * m * sh * sh)
        v0 = S.morse_cm(x, 1.146, 1943.79, 13.71, m)
        c.update(np_=[n], sh=[sh], vpot=S.morse_cm(x, 1.211, 1629.23, 17.74, m),
                 u=S.ground_state(v0, shm), v=np.zeros(n))
    return c


def grid_nd(np_, seed):
    rng = np.random.default_rng(seed)
    axes = [np.linspace(-2.5, 2.5, n) for n in np_]
    mesh = np.meshgrid(*axes[::-1], indexing="ij")[::-1]   # x fastest
    vpot = sum(0.05 * (a + 1) * q ** 2 for a, q in enumerate(mesh)) + 0.01 * mesh[0] * mesh[-1]
    psi = np.exp(-sum((q - rng.uniform(-0.4, 0.4)) ** 2 / 0.4 for q in mesh))
    psi /= np.linalg.norm(psi)
    sh = [a[1] - a[0] for a in axes]
    sm = [1836.0, 918.0, 1500.0][: len(np_)]
    # Leapfrog stability needs 2 dt E_max < 2 (a.u.); take dt at 0.3 of it.
    emax = sum((np.pi / h) ** 2 / (2 * m) for h, m in zip(sh, sm)) + np.abs(vpot).max()
    return dict(np_=list(np_), sh=sh, sm=sm, vpot=vpot.ravel(), u=psi.ravel(),
                v=np.zeros(psi.size), dt=0.3 / emax * E.SHFS)


def dC(a, b):
    return max(np.abs(a["cr"] - b["cr"]).max(), np.abs(a["ci"] - b["ci"]).max())


def run_both(orc, c, dt, nst, lc, flags, mode):
    tf = (nst - 1) * dt
    print(f"case: PSODFFT np={c['np_']} dt={dt} steps={nst} lc={lc} flags={flags}", flush=True)
    g = E.psodfft(c["np_"], dt, 0.0, tf, lc, c["sh"], c["sm"], c["u"], c["v"], c["vpot"], flags)
    o = orc.psodfft(c["np_"], dt, 0.0, tf, lc, c["sh"], c["sm"], c["u"], c["v"], c["vpot"], mode)
    assert len(g["cr"]) == len(o["cr"])
    return g, o


@pytest.mark.parametrize("modename", ["REF", "EXACT"])
def test_psodfft_fp64_1d(gpu, orc, modename):
    import oracle
    flags = E.FP64 | getattr(E, modename)
    g, o = run_both(orc, morse_case(), 5e-5, 600, 1, flags, getattr(oracle, modename))
    assert dC(g, o) <= TOL64
    assert np.abs(g["u"] - o["u"]).max() <= TOL64
    assert np.abs(g["v"] - o["v"]).max() <= TOL64


@pytest.mark.parametrize("lc", [1, 2, 7])
def test_psodfft_sampling_rule(gpu, orc, lc):
    from oracle import REF
    g, o = run_both(orc, morse_case(128), 5e-5, 50, lc, E.FP64 | E.REF, REF)
    assert dC(g, o) <= TOL64


def test_psodfft_mixed_1d(gpu, orc):
    from oracle import EXACT
    g, o = run_both(orc, morse_case(), 5e-5, 2000, 1, E.MIXED | E.EXACT, EXACT)
    assert dC(g, o) <= TOLMIX


@pytest.mark.parametrize("np_", [[32, 32], [16, 8, 8]])
@pytest.mark.parametrize("modename", ["REF", "EXACT"])
def test_psodfft_lds_multid(gpu, orc, np_, modename):
    import oracle
    c = grid_nd(np_, 3)
    g, o = run_both(orc, c, c["dt"], 200, 1, E.FP64 | getattr(E, modename),
                    getattr(oracle, modename))
    assert dC(g, o) <= TOL64


@pytest.mark.parametrize("np_", [[64, 64], [4096], [16, 16, 16], [256, 32]])
@pytest.mark.parametrize("flags,tol", [(E.FP64 | E.EXACT, TOL64), (E.MIXED | E.EXACT, TOLMIX)])
def test_psodfft_multipass(gpu, orc, np_, flags, tol):
    from oracle import EXACT
    c = grid_nd(np_, 5)
    g, o = run_both(orc, c, c["dt"], 150, 3, flags, EXACT)
    assert dC(g, o) <= tol


def test_sod_plan_batched(gpu, orc):
    from oracle import EXACT
    base = grid_nd([32, 16], 9)
    n = 32 * 16
    vp = np.stack([base["vpot"] * (1 + 0.1 * k) for k in range(3)])
    u = np.stack([np.roll(base["u"], 5 * k) for k in range(3)])
    nst, dt = 40, base["dt"]
    print("case: sod plan npk=3", flush=True)
    P = E.SodPlan([32, 16], dt, 1, base["sh"], base["sm"], vp, u, None, maxsamp=nst + 1,
                  flags=E.FP64 | E.EXACT)
    P.run(15)
    P.run(nst - 15)
    for k in range(3):
        g = P.get(k)
        o = orc.psodfft([32, 16], dt, 0.0, (nst - 1) * dt, 1, base["sh"], base["sm"], u[k],
                        np.zeros(n), vp[k], EXACT)
        assert dC(g, o) <= TOL64, k
    P.close()
