"""Gaussian initial packets (INIT_COND, init_cond.f) on the GPU against the
oracle, which tests/test_oracle_init_cond.py pins to the reference build
bit for bit.  Bar: 1e-14 (device exp/sin/cos against the host's)."""
import numpy as np
import pytest

import espec_amd as E

pytestmark = pytest.mark.gpu

XI, SH, X0, RA, RK = [-3.0, -2.5], [0.05, 0.07], [0.4, -0.3], [0.5, 0.35], [2.0, -1.5]


@pytest.mark.parametrize("mode", [E.INIT_CALC, E.INIT_GETR, E.INIT_GETR2, E.INIT_GETC,
                                  E.INIT_GETC2])
@pytest.mark.parametrize("np_", [[64, 48], [256, 128]])
def test_init_cond_2d(gpu, orc, mode, np_):
    rng = np.random.default_rng(mode)
    ea = 1 if mode in (E.INIT_GETR2, E.INIT_GETC2) else 0
    u = rng.standard_normal(np_[ea]) if mode != E.INIT_CALC else None
    v = rng.standard_normal(np_[ea]) if mode in (E.INIT_GETC, E.INIT_GETC2) else None
    print(f"case: INIT_COND mode={mode} np={np_}", flush=True)
    g = E.init_cond(np_, mode, XI, SH, X0, RA, RK, u, v)
    o = orc.init_cond(np_, mode, XI, SH, X0, RA, RK, u, v)
    assert np.abs(g[0] - o[0]).max() <= 1e-14 and np.abs(g[1] - o[1]).max() <= 1e-14


@pytest.mark.parametrize("n", [1, 100, 4096])
def test_init_cond_1d(gpu, orc, n):
    g = E.init_cond([n], E.INIT_CALC, XI, SH, X0, RA, RK)
    o = orc.init_cond([n], E.INIT_CALC, XI, SH, X0, RA, RK)
    assert np.abs(g[0] - o[0]).max() <= 1e-14 and np.abs(g[1] - o[1]).max() <= 1e-14
    assert abs(np.sum(g[0] ** 2 + g[1] ** 2) - 1.0) <= 1e-13


def test_init_cond_rejects_bad_forms(gpu):
    with pytest.raises(E.EspecError):
        E.init_cond([64], E.INIT_GETR, XI, SH, X0, RA, RK, np.ones(64))
    with pytest.raises(E.EspecError):
        E.init_cond([8, 8], 7, XI, SH, X0, RA, RK)
