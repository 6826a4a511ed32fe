"""PPSOD (ppsod.f:1-548) on the device against the oracle's restatement,
which tests/test_oracle_ppsod.py pins to the reference build.

Bars: |dC(t)|, |d psi| <= 1e-10 (FP64 leapfrog, stencil sums in another
order); VEFF bit for bit (the product rounded before the difference, as
CHPOT); the energies and projections of the samples against numpy on the
dense FD matrix (relative 1e-11)."""
import numpy as np
import pytest

import espec_amd as E
from oracle import ABS_NONE, ABS_SMOOTHW, ABS_VOPTIC

pytestmark = pytest.mark.gpu

A0A, AMU = 0.5291772083, 1836.152667539


def case_1d(n=201):
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


def case_2d(np_=(48, 32)):
    axes = [np.linspace(-2.0, 2.0, k) for k in np_]
    q1, q2 = np.meshgrid(axes[0], axes[1], indexing="xy")
    vp = 0.08 * q1 ** 2 + 0.05 * q2 ** 2 + 0.01 * q1 * q2
    psi = np.exp(-((q1 - 0.2) ** 2 + (q2 - 0.1) ** 2) / 0.3)
    psi /= np.linalg.norm(psi)
    h = [a[1] - a[0] for a in axes]
    shm = [1.0 / (24 * 1836.0 * h[0] ** 2), 1.0 / (24 * 918.0 * h[1] ** 2)]
    return dict(np_=list(np_), shm=shm, vp=vp.ravel(), u=psi.ravel(), v=np.zeros(psi.size),
                dm=(0.6 * q1 - 0.3 * q2).ravel())


def pulse(efc):
    return (efc, 2e9, 0.05, 0.0, 0.06, 2.0, 0.3, 2)


def dC(a, b):
    return max(np.abs(a["cr"] - b["cr"]).max(), np.abs(a["ci"] - b["ci"]).max())


@pytest.mark.parametrize("efc", [0, 1, 3])
@pytest.mark.parametrize("absorb", [ABS_NONE, ABS_VOPTIC, ABS_SMOOTHW])
@pytest.mark.parametrize("lc", [1, 3])
def test_ppsod_1d(gpu, orc, efc, absorb, lc):
    c = case_1d()
    va = {ABS_NONE: None, ABS_VOPTIC: c["vopt"], ABS_SMOOTHW: c["mask"]}[absorb]
    dt, tf = 2e-5, 300 * 2e-5
    print(f"case: PPSOD 1-D efc={efc} absorb={absorb} lc={lc}", flush=True)
    g = E.ppsod(c["np_"], dt, 0.0, tf, lc, c["shm"], c["u"], c["v"], c["vp"], c["dm"],
                pulse(efc), absorb=absorb, vabc=va)
    o = orc.ppsod(c["np_"], dt, 0.0, tf, lc, c["shm"], c["u"], c["v"], c["vp"], c["dm"],
                  pulse(efc), absorb, va)
    assert len(g["cr"]) == len(o["cr"])
    assert dC(g, o) <= 1e-10
    assert np.abs(g["norm"] - o["norm"]).max() <= 1e-10
    assert np.abs(g["u"] - o["u"]).max() <= 1e-10 and np.abs(g["v"] - o["v"]).max() <= 1e-10
    assert np.array_equal(g["veff"], o["veff"])


def test_ppsod_2d(gpu, orc):
    c = case_2d()
    dt, tf = 5e-5, 200 * 5e-5
    g = E.ppsod(c["np_"], dt, 0.0, tf, 2, c["shm"], c["u"], c["v"], c["vp"], c["dm"], pulse(1))
    o = orc.ppsod(c["np_"], dt, 0.0, tf, 2, c["shm"], c["u"], c["v"], c["vp"], c["dm"], pulse(1))
    assert dC(g, o) <= 1e-10
    assert np.abs(g["u"] - o["u"]).max() <= 1e-10


def test_ppsod_sample_energies_and_projections(gpu):
    """et/et0 are EIGENERG of U2 plus V2 with VEFF and with VP; the
    projections are FCORREL(W, U2), FCORREL(W, V2) (ppsod.f:276-330)."""
    import setups as S
    c = case_1d(96)
    rng = np.random.default_rng(4)
    w = rng.standard_normal((3, 96))
    dt, tf = 2e-5, 20 * 2e-5
    # 21 steps (the start-up and NINT(19.5) = 20 more): the one sample after
    # the first is at the last step, whose VEFF is the one returned
    g = E.ppsod(c["np_"], dt, 0.0, tf, 21, c["shm"], c["u"], c["v"], c["vp"], c["dm"], pulse(1),
                proj=w, energies=True)
    assert len(g["cr"]) == 2
    psi = g["u"] + 1j * g["v"]
    h0 = S.fd_matrix(c["vp"], c["shm"][0])
    he = S.fd_matrix(g["veff"], c["shm"][0])
    for h, key in ((he, "et"), (h0, "et0")):
        ref = g["u"] @ h @ g["u"] + g["v"] @ h @ g["v"]
        assert abs(g[key][-1] - ref) <= 1e-11 * abs(ref)
    assert np.abs(g["proj"][-1] - np.stack([w @ g["u"], w @ g["v"]], axis=1)).max() <= 1e-12
    assert abs(g["norm"][-1] - np.vdot(psi, psi).real) <= 1e-13


def _call_fortran_ppsod(c, prtcrl, lc, dt, tf, efc, npr, w):
    """PPSOD(ABSORB, PRTVEFF, PRTEIGVC2, PRTPULS, DIM, EFC, PRTCRL, TPABSOR,
    INFO, LC, MP, MPI, MPR, MXDCT, N, NSHOT, NPR, KP, NISPG, E0, T0, TD, TP,
    OMG, SNI, KL, DT, TI, TF, TOL, NP, ND, SH, SHM, U1, U2, V1, V2, XI, XP,
    XF, DM, VPOT, VABC, WORK, VAR, LNZVC, PRTREGION, NREG, RANGE,
    PRTONLYREIM) of libespec_f.so, by reference, hidden lengths last."""
    import ctypes as C
    from espec_amd import build as B
    E._load()
    f = C.CDLL(B.FLIB).ppsod_
    f.restype = None
    n = len(c["vp"])
    mxdct = max(n, 4096)
    lnzvc = np.zeros((mxdct, 4), order="F")
    if npr:
        lnzvc[:n, 0] = w
    u1, v1 = c["u"].copy(), c["v"].copy()
    u2, v2 = np.zeros(n), np.zeros(n)
    vpot = c["vp"].copy()
    info = C.c_int(-1)
    I = lambda x: C.byref(C.c_int(x))
    D = lambda x: C.byref(C.c_double(x))
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    pl = pulse(efc)
    strs = [b".1D", {0: b".NONE", 1: b".GAUS", 3: b".SIN2"}[efc], prtcrl, b".NONE"]
    f(I(0), I(0), I(0), I(0), *[C.c_char_p(x) for x in strs], C.byref(info), I(lc), I(2),
      I(4), I(3), I(mxdct), I(n), I(1000), I(npr), I(0), I(1), D(pl[1]), D(pl[2]), D(pl[3]),
      D(pl[4]), D(pl[5]), D(pl[6]), I(pl[7]), D(dt), D(0.0), D(tf), D(1e30),
      P(np.array([n, 1, 1], dtype=np.int32)), P(np.zeros(3, dtype=np.int32)), P(np.zeros(3)),
      P(np.array(c["shm"] + [0, 0])), P(u1), P(u2), P(v1), P(v2), P(np.zeros(3)),
      P(np.zeros(3)), P(np.zeros(3)), P(c["dm"]), P(vpot), P(np.ones(n)), P(np.zeros(n)),
      P(np.zeros(4)), P(lnzvc), I(0), I(0), P(np.zeros((5, 7))), I(0),
      *[C.c_size_t(len(x)) for x in strs])
    return info.value, lnzvc, u1, u2, v1, v2, vpot


@pytest.mark.parametrize("prtcrl,lc", [(b".PARTIAL", 2), (b".YES", 3), (b".NO", 2),
                                       (b".NO", 1)])
def test_fortran_ppsod_dropin(gpu, orc, monkeypatch, capfd, prtcrl, lc):
    c = case_1d()
    dt, tf = 2e-5, 40 * 2e-5
    w = c["u"] * np.cos(np.arange(len(c["u"])) * 0.3)
    info, lnzvc, u1, u2, v1, v2, vpot = _call_fortran_ppsod(c, prtcrl, lc, dt, tf, 1, 1, w)
    out = capfd.readouterr().out
    o = orc.ppsod(c["np_"], dt, 0.0, tf, lc, c["shm"], c["u"], c["v"], c["vp"], c["dm"],
                  pulse(1))
    assert info == 0
    assert np.abs(u2 - o["u"]).max() <= 1e-10 and np.abs(v2 - o["v"]).max() <= 1e-10
    assert np.array_equal(vpot, o["veff"])
    k = len(o["cr"])
    got = lnzvc[:k, 2] + 1j * lnzvc[:k, 3]
    want = o["cr"] + 1j * o["ci"]
    if prtcrl == b".NO":
        want[0] = 0  # C(TI) is stored only when PRTCRL prints (ppsod.f:120-125)
        if lc == 1:
            want[:] = 0  # ppsod.f:276-296: MC is never reset, nothing is stored
    assert np.abs(got - want).max() <= 1e-10
    rows = [ln for ln in out.splitlines() if ln.strip()[:1].isdigit()]
    if prtcrl == b".NO":
        assert not rows
    else:
        nf = int(np.floor((tf + dt / 2 - dt) / dt + 0.5))
        nrow = k if prtcrl == b".PARTIAL" else nf + 2  # .YES: the initial row and every step
        assert len(rows) == nrow
        first = rows[0].split()
        assert abs(float(first[3]) - 1.0) < 1e-5 and abs(float(first[-1]) - 1.0) < 1e-5
