"""S2PPSOD (s2ppsod.f:1-1143) on the device against the oracle (pinned to
the reference build in tests/test_oracle_s2ppsod.py): the printed table
(rows, their times, energies, correlations and populations), INFO, the
final states and VPOT, for pulse shapes .NONE/.GAUS/.SIN2, PRTCRL .YES,
.PARTIAL and .NO, with and without .VOPTIC, in 1-D and 2-D; and the
Fortran drop-in with eSPec's argument list, whose printed table matches
the oracle's to its printed digits.  The device sums are block partials
in a fixed order, so the table agrees to rounding (1e-12), not bit for
bit; the states agree to the leapfrog's rounding."""
import ctypes as C
import re

import numpy as np
import pytest

import espec_amd as E
from test_oracle_s2ppsod import DT, LC, TF, case, printed_close, pulse, table

pytestmark = pytest.mark.gpu

CASES = [(0, 2, 0), (1, 2, 0), (3, 1, 0), (1, 2, 1), (1, 0, 0)]


@pytest.mark.parametrize("efc,prt,absorb", CASES)
def test_s2ppsod_1d_matches_oracle(gpu, orc, efc, prt, absorb):
    import oracle
    c = case()
    va = c["vopt"] if absorb else None
    args = (c["np_"], DT, 0.0, TF, LC, c["shm"], c["xi"], c["sh"], c["m"], 0.05 * c["m"],
            c["u"], c["v"], c["vp"], c["dm"], pulse(efc), prt, 1e-3)
    print(f"case: s2ppsod 1-D efc={efc} prtcrl={prt} absorb={absorb}", flush=True)
    g = E.s2ppsod(*args, absorb=E.ABSORB_VOPTIC if absorb else E.ABSORB_NONE, vabc=va)
    o = orc.s2ppsod(*args, absorb=oracle.ABS_VOPTIC if absorb else oracle.ABS_NONE, vabc=va)
    assert g["info"] == o["info"]
    assert g["tab"].shape == o["tab"].shape
    if len(o["tab"]):
        assert np.abs(g["tab"] - o["tab"]).max() <= 1e-12 * max(1.0, np.abs(o["tab"]).max())
    for k, key in (("u", "u"), ("v", "v"), ("uprev", "u1"), ("vprev", "v1")):
        assert np.abs(g[k] - o[key]).max() <= 1e-14
    # VPOT as the reference leaves it: VP - DM E(T) of the last loop T (the
    # DO variable, accumulated), the product rounded first (chpot.f:58-62)
    t = 2 * DT
    for _ in range(int((TF + DT / 2 - 2 * DT + DT) / DT) - 1):
        t = t + DT
    assert np.array_equal(g["veff"], c["vp"] - c["dm"] * E.ef(*pulse(efc), t))


def test_s2ppsod_2d_matches_oracle(gpu, orc):
    rng = np.random.default_rng(4)
    np_ = [24, 20]
    n = 24 * 20
    x = np.linspace(-1, 1, n)
    vp = np.concatenate([0.3 * x ** 2, 0.25 * (x - 0.1) ** 2 + 0.01])
    u = np.concatenate([np.exp(-x ** 2 / 0.1), 0.2 * np.exp(-(x - 0.2) ** 2 / 0.1)])
    u /= np.linalg.norm(u)
    dm = rng.uniform(-1, 1, 2 * n)
    args = (np_, 1e-3, 0.0, 40e-3, 3, [0.02, 0.03], -1.0, 0.01, 1000.0, 20.0, u, np.zeros(2 * n),
            vp, dm, (1, 1e8, 0.01, 0.0, 0.02, 2.0, 0.3, 2), 2, 1e-2)
    print("case: s2ppsod 2-D 24x20", flush=True)
    g = E.s2ppsod(*args)
    o = orc.s2ppsod(*args)
    assert g["tab"].shape == o["tab"].shape and g["info"] == o["info"]
    assert np.abs(g["tab"] - o["tab"]).max() <= 1e-12 * max(1.0, np.abs(o["tab"]).max())
    assert np.abs(g["u"] - o["u"]).max() <= 1e-14 and np.abs(g["v"] - o["v"]).max() <= 1e-14


def fortran_s2ppsod(c, efc, prtcrl, lc, tf):
    """eSPec's argument list (espec.f:1417-1423) through libespec_f.so."""
    from espec_amd import build as B
    E._load()
    f = C.CDLL(B.FLIB).s2ppsod_
    f.restype = None
    n = c["np_"][0]
    u1, v1 = c["u"].copy(), c["v"].copy()
    u2, v2 = np.zeros(2 * n), np.zeros(2 * n)
    vpot = c["vp"].copy()
    info = C.c_int(-1)
    I = lambda x: C.byref(C.c_int(x))
    D = lambda x: C.byref(C.c_double(x))
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    pl = pulse(efc)
    strs = [b".1D", {0: b".NONE", 1: b".GAUS", 3: b".SIN2"}[efc], prtcrl, b".NONE"]
    xi = np.array([c["xi"], 0, 0])
    xf = np.array([c["xi"] + (n - 1) * c["sh"], 0, 0])
    f(I(0), I(0), I(0), I(0), *[C.c_char_p(x) for x in strs], C.byref(info), I(lc), I(2), I(4),
      I(3), I(2 * n), I(n), I(1000000), I(0), I(0), I(1), I(pl[7]), D(pl[1]), D(pl[2]),
      D(pl[3]), D(pl[4]), D(pl[5]), D(pl[6]), D(DT), D(0.0), D(tf), D(1e-3), D(1.0), D(0.0),
      D(0.05 * c["m"]), P(np.array([n, 1, 1], dtype=np.int32)), P(np.zeros(3, dtype=np.int32)),
      P(np.array([c["sh"], 0, 0])), P(np.array(c["shm"] + [0, 0])), P(u1), P(u2), P(v1), P(v2),
      P(xi), P(np.zeros(3)), P(xf), P(np.array([c["m"], 0, 0])), P(c["dm"]), P(vpot),
      P(np.ones(n)), P(np.zeros(2 * n)), P(np.zeros(4)), P(np.zeros((2 * n, 4))),
      *[C.c_size_t(len(x)) for x in strs])
    return info.value, u1, u2, v1, v2, vpot


@pytest.mark.parametrize("efc,prtcrl", [(1, b".YES"), (3, b".PARTIAL"), (0, b".NO")])
def test_fortran_s2ppsod_dropin(gpu, orc, capfd, efc, prtcrl):
    c = case()
    print(f"case: S2PPSOD drop-in efc={efc} {prtcrl.decode()}", flush=True)
    capfd.readouterr()
    info, u1, u2, v1, v2, vpot = fortran_s2ppsod(c, efc, prtcrl, LC, TF)
    out = capfd.readouterr().out
    prt = {b".NO": 0, b".PARTIAL": 1, b".YES": 2}[prtcrl]
    o = orc.s2ppsod(c["np_"], DT, 0.0, TF, LC, c["shm"], c["xi"], c["sh"], c["m"], 0.05 * c["m"],
                    c["u"], c["v"], c["vp"], c["dm"], pulse(efc), prt, 1e-3)
    assert info == o["info"]
    assert " oi " in out
    got = table(out)
    if prt == 0:
        assert len(got) == 0 and "Auto-correlation" not in out
    else:
        assert "Auto-correlation function:" in out
        assert got.shape == o["tab"].shape and printed_close(got, o["tab"])
    assert np.abs(u2 - o["u"]).max() <= 1e-14 and np.abs(v2 - o["v"]).max() <= 1e-14
    assert np.abs(u1 - o["u1"]).max() <= 1e-14 and np.abs(v1 - o["v1"]).max() <= 1e-14
