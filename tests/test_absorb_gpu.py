"""Absorbing boundaries on the GPU against the FP64 oracle.

ABSORB=.TRUE. with TPABSOR .VOPTIC (imaginary potential in both V half
phases, spofft.f:65-118; added to the leapfrog sum, sodfft.f:66-72) or
.SMOOTHW (mask after the closing half phase, spofft.f:162-176; after the
leapfrog update in ESPEC_EXACT, sodfft.f:108-116, whose branch the
reference never reaches).  The oracle's absorbers are pinned against the
reference build in tests/test_oracle_absorb.py.

Bars: |dC(t)|, |d psi| <= 1e-10 for FP64, <= 1e-6 for the mixed path.
Every case also checks that the absorber removed norm, so a kernel that
ignored it would fail.
"""
import numpy as np
import pytest

import espec_amd as E

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOLMIX = 1e-6
VOPT, SMOOTH = E.ABSORB_VOPTIC, E.ABSORB_SMOOTHW


def packet_1d(n=256):
    xi, xf = 0.5 / E.A0A, 2.8 / E.A0A
    x = np.linspace(xi, xf, n)
    m = 7.49 * E.AMU
    xe = 1.146 / E.A0A
    vpot = 0.5 * 0.35 * (x - xe) ** 2
    psi = np.exp(-((x - xe - 0.05) ** 2) / (2 * 0.06 ** 2))
    psi /= np.sqrt((psi ** 2).sum())
    return dict(np_=[n], sh=[x[1] - x[0]], sm=[m], vpot=vpot, u=psi, v=np.zeros(n), x=x)


def grid_nd(np_, seed):
    rng = np.random.default_rng(seed)
    axes = [np.linspace(-2.5, 2.5, n) for n in np_]
    mesh = np.meshgrid(*axes[::-1], indexing="ij")[::-1]   # x fastest
    vpot = sum(0.05 * (a + 1) * q ** 2 for a, q in enumerate(mesh)) + 0.01 * mesh[0] * mesh[-1]
    psi = np.exp(-sum((q - rng.uniform(-0.4, 0.4)) ** 2 / 0.4 for q in mesh))
    psi /= np.linalg.norm(psi)
    sh = [a[1] - a[0] for a in axes]
    sm = [1836.0, 918.0, 1500.0][: len(np_)]
    emax = sum((np.pi / h) ** 2 / (2 * m) for h, m in zip(sh, sm)) + np.abs(vpot).max()
    r2 = sum(q ** 2 for q in mesh).ravel()
    return dict(np_=list(np_), sh=sh, sm=sm, vpot=vpot.ravel(), u=psi.ravel(),
                v=np.zeros(psi.size), dt=0.3 / emax * E.SHFS, r2=r2)


def vabc_of(kind, n, r2=None, scale=1.0):
    """An absorber that reaches the packet's core: an imaginary potential
    (hartree) or a mask in (0, 1]."""
    t = np.arange(n)
    shape = (1.0 + np.sin(0.05 * t)) if r2 is None else r2 / r2.max()
    if kind == VOPT:
        return -0.02 * scale * shape
    return 1.0 - 2e-3 * scale * shape


def dC(a, b):
    return max(np.abs(a["cr"] - b["cr"]).max(), np.abs(a["ci"] - b["ci"]).max())


def dpsi(a, b):
    return max(np.abs(a["u"] - b["u"]).max(), np.abs(a["v"] - b["v"]).max())


def spo_both(orc, c, dt, nst, lc, flags, mode, kind, vabc):
    tf = (nst - 1) * dt
    print(f"case: PSPOFFT absorb={kind} np={c['np_']} steps={nst} lc={lc} flags={flags}",
          flush=True)
    g = E.pspofft(c["np_"], dt, 0.0, tf, lc, c["sh"], c["sm"], c["u"], c["v"], c["vpot"],
                  flags, absorb=kind, vabc=vabc)
    o = orc.pspofft_abs(c["np_"], dt, 0.0, tf, lc, c["sh"], c["sm"], c["u"], c["v"],
                        c["vpot"], kind, vabc, mode)
    p = orc.pspofft(c["np_"], dt, 0.0, tf, lc, c["sh"], c["sm"], c["u"], c["v"], c["vpot"],
                    mode)
    assert len(g["cr"]) == len(o["cr"])
    assert np.sum(o["u"] ** 2 + o["v"] ** 2) < np.sum(p["u"] ** 2 + p["v"] ** 2) - 1e-7
    return g, o


def sod_both(orc, c, dt, nst, lc, flags, mode, kind, vabc):
    tf = (nst - 1) * dt
    print(f"case: PSODFFT absorb={kind} np={c['np_']} steps={nst} lc={lc} flags={flags}",
          flush=True)
    g = E.psodfft(c["np_"], dt, 0.0, tf, lc, c["sh"], c["sm"], c["u"], c["v"], c["vpot"],
                  flags, absorb=kind, vabc=vabc)
    o = orc.psodfft_abs(c["np_"], dt, 0.0, tf, lc, c["sh"], c["sm"], c["u"], c["v"],
                        c["vpot"], kind, vabc, mode)
    assert len(g["cr"]) == len(o["cr"])
    return g, o


@pytest.mark.parametrize("kind", [VOPT, SMOOTH])
@pytest.mark.parametrize("modename", ["REF", "EXACT"])
def test_pspofft_abs_1d_fp64(gpu, orc, kind, modename):
    import oracle
    c = packet_1d()
    va = vabc_of(kind, 256, scale=1.0 if kind == VOPT else 0.1)
    g, o = spo_both(orc, c, 1e-4, 2000, 7, E.FP64 | getattr(E, modename),
                    getattr(oracle, modename), kind, va)
    assert dC(g, o) <= TOL64
    assert dpsi(g, o) <= TOL64


@pytest.mark.parametrize("kind", [VOPT, SMOOTH])
def test_pspofft_abs_1d_mixed(gpu, orc, kind):
    from oracle import EXACT
    c = packet_1d(512)
    va = vabc_of(kind, 512, scale=0.3 if kind == VOPT else 0.1)
    g, o = spo_both(orc, c, 2e-4, 2000, 1, E.MIXED | E.EXACT, EXACT, kind, va)
    assert dC(g, o) <= TOLMIX


@pytest.mark.parametrize("kind", [VOPT, SMOOTH])
@pytest.mark.parametrize("np_", [[64, 64], [16, 8, 8]])
@pytest.mark.parametrize("flags,tol", [(E.FP64 | E.EXACT, TOL64), (E.MIXED | E.EXACT, TOLMIX)])
def test_pspofft_abs_lds_multid(gpu, orc, kind, np_, flags, tol):
    from oracle import EXACT
    c = grid_nd(np_, 3)
    va = vabc_of(kind, len(c["u"]), c["r2"], scale=5.0)
    g, o = spo_both(orc, c, 4 * c["dt"], 300, 3, flags, EXACT, kind, va)
    assert dC(g, o) <= tol


@pytest.mark.parametrize("kind", [VOPT, SMOOTH])
@pytest.mark.parametrize("np_", [[256, 256], [512, 64], [32, 32, 64]])
@pytest.mark.parametrize("flags,tol", [(E.FP64 | E.EXACT, TOL64), (E.MIXED | E.EXACT, TOLMIX)])
def test_pspofft_abs_multipass(gpu, orc, kind, np_, flags, tol):
    """The multi-pass kernels: .VOPTIC rides on the V phase table of any
    row-pass family, .SMOOTHW runs the LDS-stage row pass."""
    from oracle import EXACT
    c = grid_nd(np_, 4)
    va = vabc_of(kind, len(c["u"]), c["r2"], scale=20.0)
    g, o = spo_both(orc, c, 8 * c["dt"], 40, 1, flags | E.MULTIPASS, EXACT, kind, va)
    assert dC(g, o) <= tol


@pytest.mark.parametrize("kind", [VOPT, SMOOTH])
def test_spo_plan_abs_shape_and_batch(gpu, orc, kind):
    """Three 256^2 packets with their own absorbers in one multi-pass plan;
    .SMOOTHW plans keep the row pass on the LDS-stage kernel."""
    from oracle import EXACT
    cs = [grid_nd([256, 256], 10 + k) for k in range(3)]
    vas = [vabc_of(kind, cs[0]["u"].size, cs[k]["r2"], scale=10.0 * (k + 1)) for k in range(3)]
    dt, nst = 8 * cs[0]["dt"], 30
    c0 = cs[0]
    P = E.SpoPlan(c0["np_"], dt, 1, c0["sh"], c0["sm"], np.stack([c["vpot"] for c in cs]),
                  np.stack([c["u"] for c in cs]), None, maxsamp=nst + 1,
                  flags=E.MIXED | E.EXACT, absorb=kind, vabc=np.stack(vas))
    try:
        sh = P.shape()
        print(f"case: SpoPlan absorb={kind} shape={sh}", flush=True)
        assert sh["lds"] == 0
        if kind == SMOOTH:
            assert sh["rows_pen"] == 0
        else:
            assert sh["vph"] == 1
        P.run(nst)
        P.sync()
        for k, c in enumerate(cs):
            g = P.get(k)
            o = orc.pspofft_abs(c0["np_"], dt, 0.0, (nst - 1) * dt, 1, c0["sh"], c0["sm"],
                                c["u"], c["v"], c["vpot"], kind, vas[k], EXACT)
            n = len(o["cr"])
            assert dC({"cr": g["cr"][:n], "ci": g["ci"][:n]}, o) <= TOLMIX
    finally:
        P.close()


@pytest.mark.parametrize("kind", [VOPT, SMOOTH])
@pytest.mark.parametrize("modename", ["REF", "EXACT"])
def test_psodfft_abs_1d_fp64(gpu, orc, kind, modename):
    import oracle
    c = grid_nd([256], 6)
    va = vabc_of(kind, 256, c["r2"], scale=5.0)
    g, o = sod_both(orc, c, c["dt"], 300, 3, E.FP64 | getattr(E, modename),
                    getattr(oracle, modename), kind, va)
    assert dC(g, o) <= TOL64
    assert dpsi(g, o) <= TOL64


@pytest.mark.parametrize("kind", [VOPT, SMOOTH])
@pytest.mark.parametrize("np_", [[32, 32], [128, 128]])
@pytest.mark.parametrize("flags,tol", [(E.FP64 | E.EXACT, TOL64), (E.MIXED | E.EXACT, TOLMIX)])
def test_psodfft_abs_multid(gpu, orc, kind, np_, flags, tol):
    """32x32 runs LDS-resident, 128x128 on the per-axis passes."""
    from oracle import EXACT
    c = grid_nd(np_, 7)
    va = vabc_of(kind, c["u"].size, c["r2"], scale=5.0)
    g, o = sod_both(orc, c, c["dt"], 150, 3, flags, EXACT, kind, va)
    assert dC(g, o) <= tol
    p = orc.psodfft(c["np_"], c["dt"], 0.0, 149 * c["dt"], 3, c["sh"], c["sm"], c["u"], c["v"],
                    c["vpot"], EXACT)
    assert np.sum(o["u"] ** 2 + o["v"] ** 2) < np.sum(p["u"] ** 2 + p["v"] ** 2) - 1e-9


def test_psodfft_ref_smoothw_is_unabsorbed(gpu):
    """The reference's PSODFFT never applies .SMOOTHW (sodfft.f:108); with
    ESPEC_REF the device run equals the run without absorber."""
    c = grid_nd([256], 8)
    tf = 99 * c["dt"]
    a = E.psodfft(c["np_"], c["dt"], 0.0, tf, 1, c["sh"], c["sm"], c["u"], c["v"], c["vpot"],
                  E.FP64 | E.REF, absorb=SMOOTH, vabc=vabc_of(SMOOTH, 256, c["r2"], 50.0))
    b = E.psodfft(c["np_"], c["dt"], 0.0, tf, 1, c["sh"], c["sm"], c["u"], c["v"], c["vpot"],
                  E.FP64 | E.REF)
    assert np.array_equal(a["cr"], b["cr"]) and np.array_equal(a["u"], b["u"])
