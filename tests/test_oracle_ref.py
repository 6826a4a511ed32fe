"""The oracle against the reference's own Fortran, compiled from its sources.

oracle/_ref/libespec_ref.so is built by oracle/Makefile (`make ref`) from
the reference's src/*.f where they lie, plus oracle/ref_harness.f90 (C
bindings only).  These tests pin every oracle function the product is
checked against; they skip where the reference build is absent.

The reference DIGT only fills the 1-D grid (dkeft.f:283-340 leave TC
unset for .2D/.3D), so propagations are compared in 1-D here; the
multi-dimensional paths are pinned by the separable-product tests in
test_oracle_math.py.
"""
import numpy as np
import pytest

from oracle import EXACT, REF, SHFS

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
    return dict(n=n, sh=sh, m=m, vpot=vpot, u=psi, v=np.zeros(n),
                shm=[1.0 / (24 * m * sh * sh)])


def cplx_err(a, b, c, d):
    return np.abs((a + 1j * b) - (c + 1j * d)).max()


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
@pytest.mark.parametrize("isn", [1, -1])
def test_fft_1d(orc, ref, n, isn):
    rng = np.random.default_rng(n)
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    ra, rb = ref.fft(a, b, n, n, n, isn)
    oa, ob = orc.fft(a, b, n, n, n, isn, REF)
    assert cplx_err(ra, rb, oa, ob) <= 1e-13 * n


@pytest.mark.parametrize("shape", [(16, 32), (8, 4), (32, 8)])
def test_fft_multivariate_call(orc, ref, shape):
    """FFT(A,B,NTOT,N,NSPAN,ISN) over one axis of a 2-D array (fft3d.f:1)."""
    n1, n2 = shape
    rng = np.random.default_rng(n1 * n2)
    a, b = rng.standard_normal(n1 * n2), rng.standard_normal(n1 * n2)
    for n, nspan in ((n1, n1), (n2, n1 * n2)):
        ra, rb = ref.fft(a, b, n1 * n2, n, nspan, 1)
        oa, ob = orc.fft(a, b, n1 * n2, n, nspan, 1, REF)
        assert cplx_err(ra, rb, oa, ob) <= 1e-12


@pytest.mark.parametrize("m", [3, 8, 12])
@pytest.mark.parametrize("itype", [1, -1])
def test_fftnt(orc, ref, m, itype):
    rng = np.random.default_rng(m)
    n = 1 << m
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    rx, ry = ref.fftnt(x, y, m, itype)
    ox, oy = orc.fftnt(x, y, m, itype, REF)
    assert cplx_err(rx, ry, ox, oy) <= 1e-13 * n


def test_digt_1d(orc, ref, case):
    tr, br = ref.digt([case["n"]], [case["sh"]], [case["m"]])
    to, bo = orc.digt([case["n"]], [case["sh"]], [case["m"]], REF)
    assert np.array_equal(tr, to) and br == bo


def test_spo_step(orc, ref, case):
    tc, b = orc.digt([case["n"]], [case["sh"]], [case["m"]], REF)
    dt = 1e-3
    args = ([case["n"]], b, dt / SHFS, dt / (2 * SHFS), case["u"], case["v"], tc, case["vpot"])
    ru, rv = ref.spo_step(*args)
    ou, ov = orc.spo_step(*args, REF)
    assert cplx_err(ru, rv, ou, ov) <= 1e-15


@pytest.mark.parametrize("lc", [1, 7])
def test_pspofft(orc, ref, case, lc):
    c = case
    r = ref.pspofft([c["n"]], 1e-3, 0.0, 0.5, lc, [c["sh"]], [c["m"]], c["u"], c["v"], c["vpot"])
    o = orc.pspofft([c["n"]], 1e-3, 0.0, 0.5, lc, [c["sh"]], [c["m"]], c["u"], c["v"], c["vpot"],
                    REF)
    assert len(o["cr"]) == len(r["cr"]) == orc.nsteps(0.0, 0.5, 1e-3) // lc + 1
    k = len(o["cr"])
    assert cplx_err(r["cr"][:k], r["ci"][:k], o["cr"], o["ci"]) <= 1e-12
    assert cplx_err(r["u"], r["v"], o["u"], o["v"]) <= 1e-12


@pytest.mark.parametrize("lc", [1, 7])
def test_psodfft(orc, ref, case, lc):
    c = case
    r = ref.psodfft([c["n"]], 1e-3, 0.0, 0.5, lc, [c["sh"]], [c["m"]], c["u"], c["v"], c["vpot"])
    o = orc.psodfft([c["n"]], 1e-3, 0.0, 0.5, lc, [c["sh"]], [c["m"]], c["u"], c["v"], c["vpot"],
                    REF)
    k = len(o["cr"])
    assert cplx_err(r["cr"][:k], r["ci"][:k], o["cr"], o["ci"]) <= 1e-13
    assert np.all(r["cr"][k:] == 0)


def test_efft4(orc, ref, case):
    c = case
    tc, _ = orc.digt([c["n"]], [c["sh"]], [c["m"]], REF)
    er = ref.efft4([c["n"]], c["u"], c["v"], tc, c["vpot"])
    eo = orc.efft4([c["n"]], c["u"], c["v"], tc, c["vpot"], REF)
    assert abs(er - eo) <= 1e-15 * abs(er)


@pytest.mark.parametrize("np_", [[20], [8, 9], [6, 7, 8]])
def test_au(orc, ref, np_):
    rng = np.random.default_rng(len(np_))
    n = int(np.prod(np_))
    u, vp = rng.standard_normal(n), rng.standard_normal(n)
    shm = [0.3, 0.5, 0.7][: len(np_)]
    assert np.abs(ref.au(np_, shm, vp, u) - orc.au(np_, shm, vp, u)).max() <= 1e-13


@pytest.mark.parametrize("nreort", [10, 4])
def test_lnz(orc, ref, case, nreort):
    c = case
    al1, bt1, b1 = ref.lnz([c["n"]], 10, nreort, c["shm"], c["vpot"], c["u"])
    al2, bt2, b2 = orc.lnz([c["n"]], 10, nreort, c["shm"], c["vpot"], c["u"])
    # The first coefficients agree to rounding; later ones amplify the
    # last-bit differences of the two builds' sums, as Lanczos does (the
    # propagated C(t) of test_plnz stays at 1e-13).
    scale = np.abs(al1).max()
    assert np.abs(al1[:5] - al2[:5]).max() <= 1e-13 * scale
    assert np.abs(bt1[:5] - bt2[:5]).max() <= 1e-13 * scale
    assert np.abs(al1 - al2).max() <= 1e-9 * scale
    assert np.abs(bt1 - bt2).max() <= 1e-9 * scale
    assert np.abs(b1[:3] - b2[:3]).max() <= 1e-12
    assert np.abs(b1 - b2).max() <= 1e-6


@pytest.mark.parametrize("abstol", [0.0, -1.0, 1e-6, 1e-9])
def test_dstevx(orc, ref, case, abstol):
    c = case
    al, bt, _ = orc.lnz([c["n"]], 10, 10, c["shm"], c["vpot"], c["u"])
    info, w1, z1 = ref.dstevx(al, bt[:-1], abstol)
    assert info == 0
    if abstol > 0:
        # DSTEBZ bisection: the oracle reproduces the interval midpoints.
        assert np.abs(w1 - orc.tridiag_bisect(al, bt[:-1], abstol)).max() <= 1e-15
    else:
        i2, w2, z2 = orc.tridiag_eig(al, bt[:-1])
        assert i2 == 0
        assert np.abs(w1 - w2).max() <= 1e-15
        for j in range(len(al)):
            s = np.sign(z1[:, j] @ z2[:, j])
            assert np.abs(z1[:, j] - s * z2[:, j]).max() <= 1e-12


@pytest.mark.parametrize("abstol,nsteps,tol", [(0.0, 200, 1e-13), (1e-6, 5, 1e-9)])
@pytest.mark.parametrize("lc", [1, 3])
def test_plnz(orc, ref, case, abstol, nsteps, tol, lc):
    """ABSTOL <= 0 is continuous in the Krylov data; with ABSTOL > 0 the
    bisection midpoints can move by up to ABSTOL between two builds whose
    tridiagonals differ in the last bit, so only a few steps are compared."""
    c = case
    dt = 1e-3
    tf = (nsteps - 1) * dt
    r = ref.plnz([c["n"]], 10, abstol, dt, 0.0, tf, lc, c["shm"], c["u"], c["v"], c["vpot"])
    o = orc.plnz([c["n"]], 10, abstol, dt, 0.0, tf, lc, c["shm"], c["u"], c["v"], c["vpot"], REF)
    assert r["info"] == 0
    assert len(o["cr"]) == len(r["cr"])
    assert cplx_err(r["cr"], r["ci"], o["cr"], o["ci"]) <= tol


def test_hwhm(orc, ref):
    rng = np.random.default_rng(5)
    mf = 10
    n = 1 << mf
    u, v = rng.standard_normal(n), rng.standard_normal(n)
    ru, rv = ref.hwhm(mf, 0.12, 61.44, 0.087, -0.12, u, v)
    ou, ov = orc.hwhm(mf, 0.12, 61.44, 0.087, -0.12, u, v)
    assert cplx_err(ru, rv, ou, ov) <= 1e-15


def test_ran(orc, ref):
    r, sr = ref.ran(1, 2000)
    o, so = orc.ran(1, 2000)
    assert np.array_equal(r, o) and sr == so
