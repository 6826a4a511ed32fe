"""Gaussian initial packets (INIT_COND, init_cond.f; GET_GAUSS, gauss.f)
in the oracle, against the reference build (oracle/_ref; skipped where
absent) for every INIEIGVC form, and against their closed forms: unit
norm, the phase gradient k of the plane-wave factor, the peak at x0 and
the half width rA at half maximum (GET_GAUSS = exp(-ln2 (x/rA)^2))."""
import numpy as np
import pytest

from oracle import INIT_CALC, INIT_GETC, INIT_GETC2, INIT_GETR, INIT_GETR2

XI, SH, X0, RA, RK = [-3.0, -2.5], [0.05, 0.07], [0.4, -0.3], [0.5, 0.35], [2.0, -1.5]


def eig1d(n, seed, cplx):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal(n)
    v = rng.standard_normal(n) if cplx else None
    return u, v


@pytest.mark.parametrize("mode", [INIT_CALC, INIT_GETR, INIT_GETR2, INIT_GETC, INIT_GETC2])
def test_init_cond_matches_reference(orc, ref, mode):
    np_ = [48, 40]
    eig_axis = 1 if mode in (INIT_GETR2, INIT_GETC2) else 0
    u, v = eig1d(np_[eig_axis], mode, mode in (INIT_GETC, INIT_GETC2))
    if mode == INIT_CALC:
        u = v = None
    a = orc.init_cond(np_, mode, XI, SH, X0, RA, RK, u, v)
    b = ref.init_cond(np_, mode, XI, SH, X0, RA, RK, u, v)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_init_cond_1d_matches_reference(orc, ref):
    a = orc.init_cond([300], INIT_CALC, XI, SH, X0, RA, RK)
    b = ref.init_cond([300], INIT_CALC, XI, SH, X0, RA, RK)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_init_cond_closed_forms(orc):
    n = 400
    u, v = orc.init_cond([n], INIT_CALC, XI, SH, X0, RA, RK)
    z = u + 1j * v
    assert abs(np.sum(np.abs(z) ** 2) - 1.0) < 1e-14
    x = XI[0] + SH[0] * np.arange(n)
    ph = np.unwrap(np.angle(z))
    c = np.abs(x - X0[0]) < 1.0
    assert np.allclose(np.gradient(ph[c], x[c]), RK[0], atol=1e-9)
    g = np.abs(z)
    i0 = np.argmax(g)
    assert abs(x[i0] - X0[0]) <= SH[0] / 2 + 1e-12
    # |z| ~ exp(-ln2 ((x-x0)/rA)^2): half of the maximum at x0 +- rA
    shape = np.exp(-0.69314718 * ((x - X0[0]) / RA[0]) ** 2)
    assert np.allclose(g / g[i0], shape / shape[i0], rtol=1e-12, atol=1e-300)
    # 2-D .CALC: U + iV = conj(a(x1) b(x2)) for the 1-D packets a, b, i.e.
    # the reference's 2-D packet carries the momenta -k1, -k2
    u2, v2 = orc.init_cond([n, 50], INIT_CALC, XI, SH, X0, RA, RK)
    ua, va = orc.init_cond([n], INIT_CALC, XI[:1], SH[:1], X0[:1], RA[:1], RK[:1])
    ub, vb = orc.init_cond([50], INIT_CALC, XI[1:], SH[1:], X0[1:], RA[1:], RK[1:])
    want = np.outer(ub + 1j * vb, ua + 1j * va)   # [axis 2][axis 1]
    assert np.allclose((u2 + 1j * v2).reshape(50, n), np.conj(want), rtol=0, atol=1e-15)
