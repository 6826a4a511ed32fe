"""The finer-grained routines on the device against the oracle
(oracle/espec_oracle.c, pinned to the reference build in
tests/test_oracle_ref.py and tests/test_oracle_absorb.py):

  PSPOFFTAUX (spofft.f:3-203)   one split-operator step, every absorber
  PSODFFTAUX (sodfft.f:3-129)   one leapfrog step, every absorber
  REORT      (reort.f:1-83)     Gram-Schmidt pass, projection on MFMA
  FFTNT      (cfft.f:6-106)     radix-2 transform, both directions
  PLNZ with PLNZAUX's .SMOOTHW mask (sil.f:171-180), in one call and
  through a batched plan

and the Fortran symbols of libespec_f.so for them, called with the
reference's argument lists.  Bars: 1e-12 (absolute, unit-norm states) for
single FP64 steps and transforms, 1e-10 for multi-step FP64 runs, 1e-6
for the mixed path.
"""
import ctypes as C

import numpy as np
import pytest

import espec_amd as E
from oracle import ABS_NONE, ABS_SMOOTHW, ABS_VOPTIC, EXACT, REF

pytestmark = pytest.mark.gpu

A0A, AMU = 0.5291772083, 1836.152667539


def packet(n=256):
    xi, xf = 0.5 / A0A, 2.8 / A0A
    sh = (xf - xi) / (n - 1)
    x = xi + sh * np.arange(n)
    m = 7.49 * AMU
    xe = 1.146 / A0A
    vpot = 0.0125 * (x - xe) ** 2
    psi = np.exp(-((x - xe - 0.1) ** 2) / (2 * 0.05 ** 2))
    psi /= np.sqrt((psi ** 2).sum())
    rng = np.random.default_rng(3)
    v = 0.05 * psi * rng.standard_normal(n)
    d = np.clip(np.maximum(0.3 - (x - xi) / (xf - xi), (x - xi) / (xf - xi) - 0.7), 0, None)
    return dict(n=n, sh=sh, m=m, vpot=vpot, u=psi, v=v, vopt=-0.8 * (d / 0.3) ** 2,
                mask=np.exp(-3.0 * d ** 2))


def vabc_of(p, absorb):
    return {ABS_NONE: None, ABS_VOPTIC: p["vopt"], ABS_SMOOTHW: p["mask"]}[absorb]


@pytest.mark.parametrize("mode", [REF, EXACT])
@pytest.mark.parametrize("absorb", [ABS_NONE, ABS_VOPTIC, ABS_SMOOTHW])
@pytest.mark.parametrize("shape", [[256], [32, 16]])
def test_pspofftaux(gpu, orc, mode, absorb, shape):
    p = packet(int(np.prod(shape)))
    sm = [p["m"]] * len(shape)
    sh = [p["sh"]] * len(shape)
    tc, b = orc.digt(shape, sh, sm, mode)
    dt = 1e-3
    cst1, cst2 = dt / E.SHFS, dt / (2 * E.SHFS)
    va = vabc_of(p, absorb)
    flags = E.REF if mode == REF else E.EXACT
    print(f"case: PSPOFFTAUX shape={shape} absorb={absorb} mode={mode}", flush=True)
    g = E.pspofftaux(shape, b, cst1, cst2, tc, p["u"], p["v"], p["vpot"], flags, absorb, va)
    if absorb == ABS_NONE:
        o = orc.spo_step(shape, b, cst1, cst2, p["u"], p["v"], tc, p["vpot"], mode)
    else:
        o = orc.spo_step_abs(shape, b, cst1, cst2, p["u"], p["v"], tc, p["vpot"], absorb, va,
                             mode)
    assert max(np.abs(g[0] - o[0]).max(), np.abs(g[1] - o[1]).max()) <= 1e-12


@pytest.mark.parametrize("mode", [REF, EXACT])
@pytest.mark.parametrize("absorb", [ABS_NONE, ABS_VOPTIC, ABS_SMOOTHW])
def test_psodfftaux(gpu, orc, mode, absorb):
    p = packet(256)
    tc, b = orc.digt([256], [p["sh"]], [p["m"]], mode)
    cst = 2e-4 / E.SHFS
    va = vabc_of(p, absorb)
    rng = np.random.default_rng(5)
    u1, v1 = p["u"] * (1 + 1e-3 * rng.standard_normal(256)), p["v"].copy()
    flags = E.REF if mode == REF else E.EXACT
    print(f"case: PSODFFTAUX absorb={absorb} mode={mode}", flush=True)
    g = E.psodfftaux([256], b, cst, tc, u1, p["u"], v1, p["v"], p["vpot"], flags, absorb, va)
    o = orc.sod_step_abs([256], b, cst, u1, p["u"], v1, p["v"], tc, p["vpot"], absorb,
                         va if va is not None else np.zeros(256), mode)
    for a, c in zip(g, o):
        assert np.abs(a - c).max() <= 1e-12


@pytest.mark.parametrize("n,m", [(1, 1), (37, 3), (4096, 16), (65536, 33), (201, 300)])
def test_reort(gpu, orc, n, m):
    """The MFMA projection covers ragged point counts (n not a multiple of
    4), window widths not a multiple of 16 and windows wider than n."""
    rng = np.random.default_rng(n + m)
    basis = rng.standard_normal((m, n)) / np.sqrt(n)
    w = rng.standard_normal(n)
    print(f"case: REORT n={n} m={m}", flush=True)
    g = E.reort(basis, w)
    o = orc.reort(basis, w)
    scale = np.abs(basis).max() * np.abs(w).sum() * m
    assert np.abs(g - o).max() <= 1e-13 * max(scale, 1.0)


def test_reort_orthogonalises(gpu):
    rng = np.random.default_rng(9)
    q, _ = np.linalg.qr(rng.standard_normal((5000, 24)))
    w = rng.standard_normal(5000)
    g = E.reort(q.T.copy(), w)
    assert np.abs(q.T @ g).max() <= 1e-12 * np.linalg.norm(w)


@pytest.mark.parametrize("mode", [REF, EXACT])
@pytest.mark.parametrize("m", [1, 4, 10, 15])
@pytest.mark.parametrize("itype", [1, -1])
def test_fftnt(gpu, orc, mode, m, itype):
    rng = np.random.default_rng(m)
    x, y = rng.standard_normal(1 << m), rng.standard_normal(1 << m)
    g = E.fftnt(x, y, m, itype, E.REF if mode == REF else E.EXACT)
    o = orc.fftnt(x, y, m, itype, mode)
    tol = 1e-13 * (1 << m) if itype == 1 else 1e-13
    assert max(np.abs(g[0] - o[0]).max(), np.abs(g[1] - o[1]).max()) <= tol


@pytest.mark.parametrize("flags,tol", [(E.FP64 | E.REF, 1e-10), (E.FP64 | E.EXACT, 1e-10),
                                       (E.MIXED | E.EXACT, 1e-6)])
def test_plnz_smoothw(gpu, orc, flags, tol):
    p = packet(256)
    shm = [1.0 / (24.0 * p["m"] * p["sh"] ** 2)]
    dt, nst = 2e-3, 120
    mode = REF if flags & E.REF else EXACT
    print(f"case: PLNZ .SMOOTHW flags={flags}", flush=True)
    g = E.plnz([256], 10, 0.0, dt, 0.0, (nst - 1) * dt, 1, shm, p["u"], p["v"], p["vpot"],
               flags, E.ABSORB_SMOOTHW, p["mask"])
    o = orc.plnz([256], 10, 0.0, dt, 0.0, (nst - 1) * dt, 1, shm, p["u"], p["v"], p["vpot"],
                 mode, mask=p["mask"])
    assert len(g["cr"]) == len(o["cr"])
    d = max(np.abs(g["cr"] - o["cr"]).max(), np.abs(g["ci"] - o["ci"]).max())
    assert d <= tol
    # .VOPTIC leaves PLNZAUX unchanged (sil.f:171 tests .SMOOTHW only)
    if flags == E.FP64 | E.EXACT:
        a = E.plnz([256], 10, 0.0, dt, 0.0, (nst - 1) * dt, 1, shm, p["u"], p["v"], p["vpot"],
                   flags, E.ABSORB_VOPTIC, p["vopt"])
        b = E.plnz([256], 10, 0.0, dt, 0.0, (nst - 1) * dt, 1, shm, p["u"], p["v"], p["vpot"],
                   flags)
        assert np.array_equal(a["cr"], b["cr"])


def test_sil_plan_smoothw_batched(gpu, orc):
    """Three packets, each with its own mask, in one plan run in two pieces."""
    p = packet(128)
    shm = [1.0 / (24.0 * p["m"] * p["sh"] ** 2)]
    masks = np.stack([p["mask"], np.roll(p["mask"], 7), np.ones(128)])
    u = np.stack([p["u"], np.roll(p["u"], 3), p["u"]])
    dt, nst = 2e-3, 40
    P = E.SilPlan([128], 8, 0.0, dt, 1, shm, np.stack([p["vpot"]] * 3), u, None,
                  maxsamp=nst + 1,
                  flags=E.FP64 | E.EXACT)
    P.set_absorb(E.ABSORB_SMOOTHW, masks)
    P.run(15)
    P.run(nst - 15)
    for k in range(3):
        g = P.get(k)
        o = orc.plnz([128], 8, 0.0, dt, 0.0, (nst - 1) * dt, 1, shm, u[k], np.zeros(128),
                     p["vpot"], EXACT, mask=masks[k])
        assert max(np.abs(g["cr"] - o["cr"]).max(), np.abs(g["ci"] - o["ci"]).max()) <= 1e-10
    P.close()


def _flib():
    from espec_amd import build as B
    E._load()
    return C.CDLL(B.FLIB)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def test_fortran_reort_with_leading_dimension(gpu, orc):
    """REORT(N, M, LMTREORT, MXDCT, WORK, EIGVC) with MXDCT > N."""
    n, m, mxdct = 300, 12, 512
    rng = np.random.default_rng(2)
    eig = np.asfortranarray(rng.standard_normal((mxdct, m)))
    w = rng.standard_normal(n)
    o = orc.reort(np.ascontiguousarray(eig[:n].T), w)
    f = _flib().reort_
    f.restype = None
    wf = w.copy()
    f(C.byref(C.c_int(n)), C.byref(C.c_int(m)), C.byref(C.c_int(m)), C.byref(C.c_int(mxdct)),
      _p(wf), _p(eig))
    assert np.abs(wf - o).max() <= 1e-12


def test_fortran_fftnt_and_pspofftaux(gpu, orc, monkeypatch):
    monkeypatch.setenv("ESPEC_MODE", "REF")
    lib = _flib()
    rng = np.random.default_rng(4)
    x, y = rng.standard_normal(1024), rng.standard_normal(1024)
    o = orc.fftnt(x, y, 10, -1, REF)
    f = lib.fftnt_
    f.restype = None
    f(_p(x), _p(y), C.byref(C.c_int(1024)), C.byref(C.c_int(10)), C.byref(C.c_int(-1)))
    assert max(np.abs(x - o[0]).max(), np.abs(y - o[1]).max()) <= 1e-13
    # PSPOFFTAUX(ABSORB, TPABSOR, NDIM, N, B, B1, B2, CST1, CST2, NP, U, V, TC, VPOT, VABC)
    p = packet(256)
    tc, b = orc.digt([256], [p["sh"]], [p["m"]], REF)
    cst1, cst2 = 1e-3 / E.SHFS, 1e-3 / (2 * E.SHFS)
    u, v = p["u"].copy(), p["v"].copy()
    tp = b".SMOOTHW"
    g = lib.pspofftaux_
    g.restype = None
    g(C.byref(C.c_int(1)), C.c_char_p(tp), C.byref(C.c_int(1)), C.byref(C.c_int(256)),
      C.byref(C.c_double(b)), C.byref(C.c_double(0.0)), C.byref(C.c_double(0.0)),
      C.byref(C.c_double(cst1)), C.byref(C.c_double(cst2)),
      _p(np.array([256, 1, 1], dtype=np.int32)), _p(u), _p(v), _p(np.ascontiguousarray(tc)),
      _p(p["vpot"]), _p(p["mask"]), C.c_size_t(len(tp)))
    o = orc.spo_step_abs([256], b, cst1, cst2, p["u"], p["v"], tc, p["vpot"], ABS_SMOOTHW,
                         p["mask"], REF)
    assert max(np.abs(u - o[0]).max(), np.abs(v - o[1]).max()) <= 1e-12


def test_fortran_plnzaux_one_step(gpu, orc, monkeypatch):
    """PLNZAUX(ABSORB, CHANGE, DIM, TPABSOR, INFO, LMTREORT, MP, MXDCT, N,
    ABSTOL, CST, NP, IWORK, SHM, VPOT, VABC, U, V, WU, WV, WP, WORK, VAR,
    LNZVC): one step with the .SMOOTHW mask, against one oracle step."""
    monkeypatch.setenv("ESPEC_MODE", "REF")
    lib = _flib()
    p = packet(256)
    shm = np.array([1.0 / (24.0 * p["m"] * p["sh"] ** 2), 0, 0])
    cst = 2e-3 / E.SHFS
    n, mp = 256, 10
    u, v = p["u"].copy(), p["v"].copy()
    work = [np.zeros(n) for _ in range(4)]
    lnzvc = np.zeros((n, mp + 2), order="F")
    info = C.c_int(-1)
    f = lib.plnzaux_
    f.restype = None
    dim, tp = b".1D", b".SMOOTHW"
    f(C.byref(C.c_int(1)), C.byref(C.c_int(0)), C.c_char_p(dim), C.c_char_p(tp), C.byref(info),
      C.byref(C.c_int(300)), C.byref(C.c_int(mp)), C.byref(C.c_int(n)), C.byref(C.c_int(n)),
      C.byref(C.c_double(0.0)), C.byref(C.c_double(cst)),
      _p(np.array([n, 1, 1], dtype=np.int32)), _p(np.zeros(64, dtype=np.int32)), _p(shm),
      _p(p["vpot"]), _p(p["mask"]), _p(u), _p(v), *[_p(a) for a in work], _p(np.zeros(4)),
      _p(lnzvc), C.c_size_t(len(dim)), C.c_size_t(len(tp)))
    assert info.value == 0
    o = orc.plnz([n], mp, 0.0, 2e-3, 0.0, 0.0, 1, shm[:1], p["u"], p["v"], p["vpot"], REF,
                 mask=p["mask"])
    assert max(np.abs(u - o["u"]).max(), np.abs(v - o["v"]).max()) <= 1e-12
