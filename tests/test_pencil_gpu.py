"""Multi-pass split-operator plans with pinned launch shapes: the stage
kernels and the register-pencil kernels, with and without the grid-stride
loop and prefetch, against the FP64 oracle (|dC| <= 1e-10 FP64, <= 1e-6
mixed)."""
import numpy as np
import pytest

import espec_amd as E
from espec_amd import configs

pytestmark = pytest.mark.gpu

SHAPES = {
    "stage": dict(rows_pen=0, tile_pen=0, grid=0),
    "stage_loop": dict(rows_pen=0, tile_pen=0, grid=16),
    "pen": dict(rows_pen=1, tile_pen=1, pen_pf=0, grid=0, xmap=0),
    "pen_rows": dict(rows_pen=1, tile_pen=0, pen_pf=0, grid=0),
    "pen_cols": dict(rows_pen=0, tile_pen=1, pen_pf=0, grid=0, xmap=1),
    "pen_loop_pf": dict(rows_pen=1, tile_pen=1, pen_pf=1, grid=24, xmap=1),
    "pen_loop": dict(rows_pen=1, tile_pen=1, pen_pf=0, grid=40, xmap=0),
    "pen_tile": dict(rows_pen=1, tile_pen=2, pen_pf=0, grid=0, xmap=1),
    "pen_tile_nox": dict(rows_pen=0, tile_pen=2, pen_pf=0, grid=0, xmap=0),
    "wide": dict(rows_pen=2, tile_pen=3, pen_pf=0, grid=0, xmap=1),
    "wide_rows": dict(rows_pen=2, tile_pen=1, pen_pf=1, grid=0, xmap=1),
    "wide_cols": dict(rows_pen=1, tile_pen=3, pen_pf=0, grid=0, xmap=0),
}
# FP32-transform row pass: complex64 storage only
SHAPES_MIXED = {
    "pen_sp": dict(rows_pen=3, tile_pen=1, pen_pf=0, grid=0, xmap=1),
    "pen_sp_loop": dict(rows_pen=3, tile_pen=1, pen_pf=1, grid=24, xmap=1),
}


def pinned(plan, name, np0):
    kw = dict(SHAPES.get(name) or SHAPES_MIXED[name])
    if kw.get("rows_pen") in (1, 3):
        kw.update(rows=1024 // np0, rows_nt=64)
    elif kw.get("rows_pen") == 2:
        kw.update(rows=256 // np0, rows_nt=64)
    plan.set_shape(**kw)
    got = plan.shape()
    for k, v in kw.items():
        assert got[k] == v, (k, got)


def dC(g, o):
    k = min(len(g["cr"]), len(o["cr"]))
    return max(np.abs(g["cr"][:k] - o["cr"][:k]).max(), np.abs(g["ci"][:k] - o["ci"][:k]).max())


@pytest.mark.parametrize("shape", list(SHAPES))
@pytest.mark.parametrize("flags,tol", [(E.FP64, 1e-10), (E.MIXED, 1e-6)])
def test_config2_shapes(gpu, orc, shape, flags, tol):
    from oracle import EXACT
    c = configs.config2()
    nst, lc = 40, 1
    print(f"case: config2 shape={shape} flags={flags}", flush=True)
    plan = E.SpoPlan(c.np, c.dt_fs, lc, c.sh, c.sm, c.vpot[None], c.u[None], c.v[None],
                     maxsamp=nst + 1, flags=flags | E.EXACT)
    pinned(plan, shape, c.np[0])
    plan.run(nst)
    g = plan.get(0)
    o = orc.pspofft(c.np, c.dt_fs, 0.0, (nst - 1) * c.dt_fs, lc, c.sh, c.sm, c.u, c.v, c.vpot,
                    EXACT)
    assert dC(g, o) <= tol
    plan.close()


@pytest.mark.parametrize("shape", list(SHAPES_MIXED))
def test_config2_fp32_row_transforms(gpu, orc, shape):
    from oracle import EXACT
    c = configs.config2()
    nst = 40
    print(f"case: config2 mixed shape={shape}", flush=True)
    plan = E.SpoPlan(c.np, c.dt_fs, 1, c.sh, c.sm, c.vpot[None], c.u[None], c.v[None],
                     maxsamp=nst + 1, flags=E.MIXED | E.EXACT)
    pinned(plan, shape, c.np[0])
    plan.run(nst)
    g = plan.get(0)
    o = orc.pspofft(c.np, c.dt_fs, 0.0, (nst - 1) * c.dt_fs, 1, c.sh, c.sm, c.u, c.v, c.vpot,
                    EXACT)
    assert dC(g, o) <= 1e-6
    plan.close()


def test_fp32_row_transforms_refused_for_fp64(gpu):
    c = configs.config2()
    plan = E.SpoPlan(c.np, c.dt_fs, 1, c.sh, c.sm, c.vpot[None], c.u[None], c.v[None],
                     maxsamp=4, flags=E.FP64 | E.EXACT)
    with pytest.raises(E.EspecError):
        plan.set_shape(rows_pen=3, rows=1024 // c.np[0], rows_nt=64)
    plan.close()


@pytest.mark.parametrize("shape", ["pen", "pen_tile", "wide", "pen_sp"])
def test_config2_phase_from_v(gpu, orc, shape):
    """Mixed plans above 64 MB form exp(-i cst2 V) in the row pass from V
    (table-based FP64 sincos) instead of reading a phase table: 80 copies
    of the 256x256 packet, packets 0 and 79 against the oracle."""
    from oracle import EXACT
    c = configs.config2()
    nst, npk = 40, 80
    print(f"case: config2 x{npk} phase from V shape={shape}", flush=True)
    plan = E.SpoPlan(c.np, c.dt_fs, 1, c.sh, c.sm, np.repeat(c.vpot[None], npk, axis=0),
                     np.repeat(c.u[None], npk, axis=0), None, maxsamp=nst + 1,
                     flags=E.MIXED | E.EXACT)
    pinned(plan, shape, c.np[0])
    assert plan.shape()["vph"] == 0
    plan.run(nst)
    o = orc.pspofft(c.np, c.dt_fs, 0.0, (nst - 1) * c.dt_fs, 1, c.sh, c.sm, c.u, c.v, c.vpot,
                    EXACT)
    for k in (0, npk - 1):
        assert dC(plan.get(k, psi=False), o) <= 1e-6, k
    plan.close()


@pytest.mark.parametrize("shape", ["pen", "pen_loop_pf", "stage", "pen_tile", "wide", "wide_cols"])
def test_3d_shapes(gpu, orc, shape):
    """64^3 (every axis a pencil length), two runs continuing one propagation."""
    from oracle import EXACT
    c = configs.config3(n=64)
    nst, lc = 12, 2
    print(f"case: 3d 64^3 shape={shape}", flush=True)
    plan = E.SpoPlan(c.np, c.dt_fs, lc, c.sh, c.sm, c.vpot[None], c.u[None], c.v[None],
                     maxsamp=nst // lc + 1, flags=E.FP64 | E.EXACT)
    pinned(plan, shape, c.np[0])
    plan.run(5)
    plan.run(nst - 5)
    g = plan.get(0)
    o = orc.pspofft(c.np, c.dt_fs, 0.0, (nst - 1) * c.dt_fs, lc, c.sh, c.sm, c.u, c.v, c.vpot,
                    EXACT)
    assert dC(g, o) <= 1e-10


@pytest.mark.parametrize("shape", ["pen", "pen_loop_pf", "pen_loop", "pen_tile", "wide"])
@pytest.mark.parametrize("flags,tol", [(E.FP64, 1e-10), (E.MIXED, 1e-6)])
def test_batch_per_packet_masses(gpu, orc, shape, flags, tol):
    """Five 128^2 packets with their own masses, potentials and states;
    C(t) every third step."""
    from oracle import EXACT
    rng = np.random.default_rng(21)
    pk = [configs.packet2d(128, rng.uniform(0.1, 0.2, 2), rng.uniform(0.8, 1.4, 2),
                           rng.uniform(500, 2000, 2), rng.uniform(0.1, 0.5, 2),
                           rng.uniform(0.1, 0.3, 2), 0.5, 30) for _ in range(5)]
    nst, lc = 30, 3
    n = 128 * 128
    print(f"case: batch 5x128^2 shape={shape} flags={flags}", flush=True)
    plan = E.SpoPlan([128, 128], pk[0].dt_fs, lc, pk[0].sh,
                     np.array([p.sm for p in pk], dtype=np.float64),
                     np.stack([p.vpot.reshape(n) for p in pk]),
                     np.stack([p.u.reshape(n) for p in pk]),
                     np.stack([p.v.reshape(n) for p in pk]),
                     maxsamp=nst // lc + 1, flags=flags | E.EXACT)
    pinned(plan, shape, 128)
    plan.run(nst)
    for k, p in enumerate(pk):
        o = orc.pspofft(p.np, p.dt_fs, 0.0, (nst - 1) * p.dt_fs, lc, p.sh, p.sm, p.u, p.v,
                        p.vpot, EXACT)
        assert dC(plan.get(k), o) <= tol, k
    plan.close()


def test_shape_change_mid_run_keeps_the_state(gpu):
    """A plan's step graphs are replayed once when they are captured, and
    psi restored from psi0, only while the plan is at step 0: pinning a
    shape after 7 steps must continue the same propagation (bit for bit
    against a plan pinned to that shape from the start)."""
    c = configs.config2()
    nst = 20
    print("case: config2 mixed shape change mid-run", flush=True)
    mk = lambda: E.SpoPlan(c.np, c.dt_fs, 1, c.sh, c.sm, c.vpot[None], c.u[None], c.v[None],
                           maxsamp=nst + 1, flags=E.MIXED | E.EXACT)
    a, b = mk(), mk()
    pinned(a, "pen_tile", c.np[0])
    a.run(7)
    a.run(nst - 7)
    pinned(b, "pen_tile", c.np[0])
    b.run(7)
    pinned(b, "pen_tile", c.np[0])  # captures the step graphs again, at step 7
    b.run(nst - 7)
    ga, gb = a.get(0), b.get(0)
    for key in ("cr", "ci", "u", "v"):
        assert np.array_equal(ga[key], gb[key]), key
    a.close()
    b.close()
