"""Device entry points other than the propagators, against the oracle.

FFT, AU, FCORREL, LNZ, DSTEVX, LANCZS and the correlation -> spectrum
chain run through libespec_hip.so; each is compared with the oracle on the
same inputs (FP64: integer/index work bit-exact, arithmetic to rounding).
"""
import numpy as np
import pytest

import espec_amd as E
import setups as S

pytestmark = pytest.mark.gpu


def cerr(a, b, c, d):
    return np.abs((a + 1j * b) - (c + 1j * d)).max()


@pytest.mark.parametrize("n", [2, 8, 64, 256, 1024, 4096])
@pytest.mark.parametrize("isn", [1, -1])
@pytest.mark.parametrize("mode", ["REF", "EXACT"])
def test_fft_1d(gpu, orc, n, isn, mode):
    import oracle
    rng = np.random.default_rng(n)
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    print(f"case: fft n={n} isn={isn} {mode}", flush=True)
    ga, gb = E.fft(a, b, n, n, n, isn, getattr(E, mode))
    oa, ob = orc.fft(a, b, n, n, n, isn, getattr(oracle, mode))
    assert cerr(ga, gb, oa, ob) <= 1e-13 * n


@pytest.mark.parametrize("shape", [(16, 32), (64, 8, 4)])
def test_fft_multivariate(gpu, orc, shape):
    from oracle import REF
    rng = np.random.default_rng(len(shape))
    ntot = int(np.prod(shape))
    a, b = rng.standard_normal(ntot), rng.standard_normal(ntot)
    span = 1
    for n in shape:
        span *= n
        print(f"case: fft ntot={ntot} n={n} nspan={span}", flush=True)
        ga, gb = E.fft(a, b, ntot, n, span, 1, E.REF)
        oa, ob = orc.fft(a, b, ntot, n, span, 1, REF)
        assert cerr(ga, gb, oa, ob) <= 1e-12


@pytest.mark.parametrize("np_", [[20], [8, 9], [6, 7, 8], [256, 256]])
def test_au(gpu, orc, np_):
    rng = np.random.default_rng(len(np_))
    n = int(np.prod(np_))
    u, vp = rng.standard_normal(n), rng.standard_normal(n)
    shm = [0.3, 0.5, 0.7][: len(np_)]
    print(f"case: au {np_}", flush=True)
    assert np.abs(E.au(np_, shm, vp, u) - orc.au(np_, shm, vp, u)).max() <= 1e-12


@pytest.mark.parametrize("n", [1, 7, 1000, 1 << 20])
def test_fcorrel(gpu, orc, n):
    rng = np.random.default_rng(n)
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    print(f"case: fcorrel n={n}", flush=True)
    assert abs(E.fcorrel(a, b) - orc.fcorrel(a, b)) <= 1e-13 * max(1.0, np.sqrt(n))


def harmonic(n=256):
    m = 7.49 * E.AMU
    x, sh = S.grid(n, 0.5, 2.8)
    v = S.ohs(x, 1.146, 1943.79, m)
    shm = 1.0 / (24 * m * sh * sh)
    psi = np.exp(-((x - 1.25 / E.A0A) ** 2) / 0.01)
    return x, v, [shm], psi / np.linalg.norm(psi)


@pytest.mark.parametrize("m,nreort", [(10, 10), (12, 4), (40, 40)])
def test_lnz(gpu, orc, m, nreort):
    x, v, shm, psi = harmonic()
    print(f"case: lnz m={m} nreort={nreort}", flush=True)
    al1, bt1, b1 = E.lnz([len(x)], m, nreort, shm, v, psi)
    al2, bt2, b2 = orc.lnz([len(x)], m, nreort, shm, v, psi)
    scale = np.abs(al2).max()
    assert np.abs(al1[:5] - al2[:5]).max() <= 1e-13 * scale
    assert np.abs(bt1[:5] - bt2[:5]).max() <= 1e-13 * scale
    assert np.abs(b1[:3] - b2[:3]).max() <= 1e-12
    # Krylov invariants that do not depend on summation order
    V = b1[: min(m, nreort)].T
    assert np.abs(V.T @ V - np.eye(V.shape[1])).max() <= 1e-8


@pytest.mark.parametrize("abstol", [0.0, 1e-6, 1e-10])
@pytest.mark.parametrize("m", [5, 10, 32])
def test_dstevx(gpu, orc, abstol, m):
    rng = np.random.default_rng(m)
    d, e = rng.standard_normal(m), rng.standard_normal(m - 1)
    print(f"case: dstevx m={m} abstol={abstol}", flush=True)
    w, z = E.dstevx(d, e, abstol)
    info, w2, z2 = orc.tridiag_eig(d, e)
    if abstol > 0:
        w2 = orc.tridiag_bisect(d, e, abstol)
    assert np.abs(w - w2).max() <= 1e-14 * max(1.0, np.abs(w2).max())
    for j in range(m):
        s = np.sign(z[:, j] @ z2[:, j])
        assert np.abs(z[:, j] - s * z2[:, j]).max() <= 1e-12


@pytest.mark.parametrize("abstol", [0.0, 1e-6])
@pytest.mark.parametrize("n,nreort", [(201, 50), (201, 201), (64, 64)])
def test_lanczs(gpu, orc, n, nreort, abstol):
    m = 7.49 * 1836.152667539
    x = np.linspace(0.5, 2.8, n) / 0.5291772083
    h = x[1] - x[0]
    v = 0.5 * 0.1 * (x - 2.2) ** 2
    shm = [1.0 / (24 * m * h * h)]
    print(f"case: lanczs n={n} nreort={nreort} abstol={abstol}", flush=True)
    ev, vc = E.lanczs([n], 1, 4, nreort, shm, v, abstol)
    info, ev2, vc2 = orc.lanczs([n], 1, 4, nreort, shm, v, abstol)
    # with ABSTOL > 0 a midpoint may sit in the neighbouring interval when
    # the two tridiagonals differ in the last bit
    assert np.abs(ev - ev2).max() <= max(1e-12, abstol)
    for k in range(4):
        assert np.abs(vc[k] - vc2[k]).max() <= 1e-9


@pytest.mark.parametrize("t,mf", [(12, 12), (5, 14)])
def test_assemble_corr(gpu, orc, t, mf):
    tab = S.load(f"test{t}_table")
    mc = (1 << mf) // 2
    print(f"case: assemble test{t}", flush=True)
    u, v = E.assemble_corr(mf, tab[:mc, 2], tab[:mc, 3])
    uo, vo = orc.assemble_corr(mf, tab[:mc, 2], tab[:mc, 3])
    assert np.array_equal(u, uo) and np.array_equal(v, vo)


def test_hwhm(gpu, orc):
    rng = np.random.default_rng(5)
    mf = 12
    u, v = rng.standard_normal(1 << mf), rng.standard_normal(1 << mf)
    gu, gv = E.hwhm(mf, 0.122, 249.856, 0.087, -0.12, u, v)
    ou, ov = orc.hwhm(mf, 0.122, 249.856, 0.087, -0.12, u, v)
    # device cos/sin/exp agree with the host's to an ulp or two
    assert cerr(gu, gv, ou, ov) <= 2e-14


SPECTRA = [
    (12, 12, 1, 1e-3, None, 0.0, 0.0, "test12_corr"),
    (5, 14, 1, 1e-5, None, 0.0, 0.0, "test5_corr"),
    (4, 12, 0, 0.0, 0.087, -0.12, 19.562, "test4_corr"),
    (8, 13, 0, 0.0, 0.08, 0.0, 0.0, "correl8_corr"),
]


@pytest.mark.parametrize("mode", ["REF", "EXACT"])
@pytest.mark.parametrize("t,mf,iwin,wp,hw,de,ade,src", SPECTRA)
def test_spectrum_chain(gpu, orc, t, mf, iwin, wp, hw, de, ade, src, mode):
    """HWHM + SPECTRUMTD on the device against the oracle (to rounding) and
    against the printed spectrum of the reference's regression output."""
    import oracle
    cor, spec = S.load(src), S.load(f"test{t}_spec")
    tf = -cor[0, 0]
    dtf = tf / (1 << (mf - 1))
    u, v = cor[:, 1].copy(), cor[:, 2].copy()
    print(f"case: spectrum test{t} {mode}", flush=True)
    gu, gv, ou, ov = u, v, u, v
    if hw is not None:
        gu, gv = E.hwhm(mf, dtf, tf, hw, de, u, v)
        ou, ov = orc.hwhm(mf, dtf, tf, hw, de, u, v)
    e, w, wi = E.spectrumtd(iwin, mf, dtf, tf, wp, ade, gu, gv, getattr(E, mode))
    eo, wo, wio = orc.spectrumtd(iwin, mf, dtf, tf, wp, ade, ou, ov, getattr(oracle, mode))
    assert np.array_equal(e, eo)
    assert np.abs(w - wo).max() <= 1e-13 and np.abs(wi - wio).max() <= 1e-13
    assert np.abs(w - spec[:, 1]).max() <= 6e-8 and np.abs(wi - spec[:, 3]).max() <= 6e-8


@pytest.mark.parametrize("ti,tf,dt,mf", [(0.0, 250.0, 1e-4, 12), (0.0, 250.0, 5e-5, 14),
                                         (0.0, 150.0, 5e-5, 14), (0.0, 250.0, 2e-3, 12)])
def test_lc_rule(gpu, orc, ti, tf, dt, mf):
    assert E.lc_rule(ti, tf, dt, mf) == orc.lc_rule(ti, tf, dt, mf)


@pytest.mark.parametrize("np_", [[256], [32, 16], [8, 8, 4]])
@pytest.mark.parametrize("mode", ["REF", "EXACT"])
def test_efft4(gpu, orc, np_, mode):
    import oracle
    rng = np.random.default_rng(len(np_))
    n = int(np.prod(np_))
    u, v, vp = rng.standard_normal(n), rng.standard_normal(n), rng.uniform(0, 0.1, n)
    sh, sm = [0.05, 0.07, 0.09][: len(np_)], [1836.0, 918.0, 500.0][: len(np_)]
    tc, b = orc.digt(np_, sh, sm, getattr(oracle, mode))
    print(f"case: efft4 {np_} {mode}", flush=True)
    e = E.efft4(np_, b, u, v, tc, vp, getattr(E, mode))
    eo = orc.efft4(np_, u, v, tc, vp, getattr(oracle, mode))
    assert abs(e - eo) <= 1e-12 * abs(eo)


@pytest.mark.parametrize("np_", [[201], [16, 12], [6, 7, 8]])
def test_eigenerg_ecnorm(gpu, orc, np_):
    rng = np.random.default_rng(len(np_) + 10)
    n = int(np.prod(np_))
    u, vp = rng.standard_normal(n), rng.uniform(0, 0.1, n)
    shm = [0.3, 0.5, 0.7][: len(np_)]
    print(f"case: eigenerg {np_}", flush=True)
    assert abs(E.eigenerg(np_, shm, u, vp) - orc.eigenerg(np_, shm, u, vp)) <= 1e-12 * n
    assert abs(E.ecnorm(u) - np.sqrt(orc.fcorrel(u, u))) <= 1e-13 * np.sqrt(n)
