"""The oracle's S2PPSOD (s2ppsod.f:1-1143): two surfaces coupled by the
coordinate operator (the routine's active TPCRS = '.CORD'), each advanced
by PSODAUX (sod.f:3-122) with VEFF = VP - DM E(T).  Pinned:

- against the reference's own S2PPSOD compiled from its sources (oracle/_ref,
  skipped where that build is absent), run in a child process: the printed
  table (FORMAT 1021, compared to its printed digits) and the final
  states U1, U2, V1, V2 (to rounding), for pulse shapes .NONE, .GAUS and
  .SIN2, PRTCRL .YES and .PARTIAL, with and without .VOPTIC;
- by the mathematics: without coupling (AA = 0) each surface is the
  oracle's single-surface PPSOD (pinned to the reference build in
  tests/test_oracle_ppsod.py) on the same steps; exchanging the surfaces
  exchanges the A and B columns; the coupling alone conserves the total
  norm to the leapfrog's accuracy and moves population between the
  surfaces.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import ABS_NONE, ABS_VOPTIC

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
A0A, AMU = 0.5291772083, 1836.152667539


def case(n=201, b_amp=0.3):
    xi, xf = 0.5 / A0A, 2.8 / A0A
    x = np.linspace(xi, xf, n)
    sh = x[1] - x[0]
    m = 7.49 * AMU
    xe = 1.146 / A0A
    vpa = 0.5 * 0.35 * (x - xe) ** 2
    vpb = 0.5 * 0.30 * (x - xe - 0.1) ** 2 + 0.02
    ga = np.exp(-((x - xe - 0.05) ** 2) / (2 * 0.06 ** 2))
    gb = b_amp * np.exp(-((x - xe + 0.08) ** 2) / (2 * 0.05 ** 2))
    nrm = np.sqrt(np.sum(ga ** 2) + np.sum(gb ** 2))
    d = np.clip(np.maximum(0.3 - (x - xi) / (xf - xi), (x - xi) / (xf - xi) - 0.7), 0, None)
    return dict(np_=[n], xi=xi, sh=sh, m=m, shm=[1.0 / (24.0 * m * sh * sh)],
                u=np.concatenate([ga, gb]) / nrm, v=np.zeros(2 * n),
                vp=np.concatenate([vpa, vpb]),
                dm=np.concatenate([0.8 * (x - xe), -0.5 * (x - xe)]),
                vopt=-0.3 * (d / 0.3) ** 2)


def pulse(efc):
    return (efc, 2e9, 0.05, 0.0, 0.06, 2.0, 0.3, 2)


DT, TF, LC = 2e-5, 60 * 2e-5, 2

CHILD = r"""
import sys, numpy as np
sys.path.insert(0, {tests!r}); sys.path.insert(0, {oracle!r})
from oracle import Reference
from test_oracle_s2ppsod import case, pulse, DT, TF, LC
c = case()
va = c["vopt"] if {absorb} else None
r = Reference().s2ppsod(c["np_"], DT, 0.0, TF, LC, c["shm"], c["xi"], c["sh"], c["m"],
                        {aa} * c["m"], c["u"], c["v"], c["vp"], c["dm"], pulse({efc}), {prt},
                        1e-3, {absorb}, va)
sys.stdout.flush()
np.save({out!r}, np.concatenate([[r["info"]], r["u"], r["v"], r["u1"], r["v1"]]))
"""

ROW = re.compile(r"^\s*(\S+)" + r"\s+(\S+)" * 11 + r"\s*$")


def table(text):
    rows = []
    for ln in text.splitlines():
        m = ROW.match(ln)
        if m:
            try:
                rows.append([float(x) for x in m.groups()])
            except ValueError:
                pass
    return np.array(rows)


# printed precision per column of FORMAT 1021: F14.6, E12.6 x2, F8.5 x2,
# E10.4, F8.5 x2, E10.4 x3, F8.6; an Ew.d field rounds to half a unit of
# its d-th digit, at most 5 10^-d of a value whose mantissa is >= 0.1
ABS = np.array([5e-7, 0, 0, 5e-6, 5e-6, 0, 5e-6, 5e-6, 0, 0, 0, 5e-7])
REL = np.array([0, 5e-6, 5e-6, 0, 0, 5e-4, 0, 0, 5e-4, 5e-4, 5e-4, 0])


def printed_close(a, b):
    return np.all(np.abs(a - b) <= ABS + REL * np.abs(b) + 1e-300)


@pytest.mark.parametrize("efc,prt,absorb", [(0, 2, 0), (1, 2, 0), (3, 1, 0), (1, 2, 1)])
def test_s2ppsod_matches_reference(orc, ref, tmp_path, efc, prt, absorb):
    out = str(tmp_path / "r.npy")
    code = CHILD.format(tests=os.path.join(ROOT, "tests"), oracle=os.path.join(ROOT, "oracle"),
                        absorb=absorb, aa=0.05, efc=efc, prt=prt, out=out)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, timeout=300,
                       encoding="latin-1")
    assert r.returncode == 0, r.stderr[-2000:]
    ref_tab = table(r.stdout)
    res = np.load(out)
    c = case()
    va = c["vopt"] if absorb else None
    o = orc.s2ppsod(c["np_"], DT, 0.0, TF, LC, c["shm"], c["xi"], c["sh"], c["m"], 0.05 * c["m"],
                    c["u"], c["v"], c["vp"], c["dm"], pulse(efc), prt, 1e-3,
                    ABS_VOPTIC if absorb else ABS_NONE, va)
    n2 = len(c["u"])
    assert int(res[0]) == o["info"]
    for k, key in enumerate(("u", "v", "u1", "v1")):
        assert np.abs(res[1 + k * n2:1 + (k + 1) * n2] - o[key]).max() <= 1e-14
    assert ref_tab.shape == o["tab"].shape and len(ref_tab) > 10
    assert printed_close(o["tab"], ref_tab)


def test_without_coupling_each_surface_is_ppsod(orc):
    c = case()
    n = c["np_"][0]
    # PPSOD takes NINT((TF+DT/2-T)/DT) loop steps, S2PPSOD's DO loop
    # INT((TF+DT/2-(TI+2DT)+DT)/DT): equal for (TF-TI)/DT = 40.6
    tf = 40.6 * DT
    o = orc.s2ppsod(c["np_"], DT, 0.0, tf, 1, c["shm"], c["xi"], c["sh"], c["m"], 0.0, c["u"],
                    c["v"], c["vp"], c["dm"], pulse(1), 1, 1e30)
    for k in range(2):
        s = slice(k * n, (k + 1) * n)
        p = orc.ppsod(c["np_"], DT, 0.0, tf, 1, c["shm"], c["u"][s], c["v"][s], c["vp"][s],
                      c["dm"][s], pulse(1))
        assert np.array_equal(o["u"][s], p["u"]) and np.array_equal(o["v"][s], p["v"])
        col_r, col_i = (3, 4) if k == 0 else (6, 7)
        # rows: psi0, then every step of the loop; ppsod's samples: psi0,
        # the start-up, then every step
        assert np.abs(o["tab"][1:, col_r] - p["cr"][2:2 + len(o["tab"]) - 1]).max() <= 1e-15
        assert np.abs(o["tab"][1:, col_i] - p["ci"][2:2 + len(o["tab"]) - 1]).max() <= 1e-15


def test_exchanging_surfaces_exchanges_columns(orc):
    c = case()
    n = c["np_"][0]
    sw = lambda a: np.concatenate([a[n:], a[:n]])
    args = (c["np_"], DT, 0.0, TF, 1, c["shm"], c["xi"], c["sh"], c["m"], 0.05 * c["m"])
    a = orc.s2ppsod(*args, c["u"], c["v"], c["vp"], c["dm"], pulse(3), 1, 1e30)
    b = orc.s2ppsod(*args, sw(c["u"]), sw(c["v"]), sw(c["vp"]), sw(c["dm"]), pulse(3), 1, 1e30)
    assert np.array_equal(a["u"], sw(b["u"])) and np.array_equal(a["v"], sw(b["v"]))
    ta, tb = a["tab"][1:], b["tab"][1:]
    for ca, cb in ((3, 6), (4, 7), (5, 8), (9, 10)):
        assert np.array_equal(ta[:, ca], tb[:, cb]) and np.array_equal(ta[:, cb], tb[:, ca])


def test_coupling_moves_population_and_keeps_the_norm(orc):
    c = case(b_amp=0.0)
    o = orc.s2ppsod(c["np_"], DT, 0.0, 400 * DT, 1, c["shm"], c["xi"], c["sh"], c["m"],
                    0.05 * c["m"], c["u"], c["v"], c["vp"], c["dm"], pulse(0), 1, 1e-4)
    t = o["tab"]
    assert t[0, 10] == 0.0 and t[-1, 10] > 1e-5  # population of surface B
    assert np.abs(t[1:, 11] - 1.0).max() <= 1e-6 and o["info"] == 0
