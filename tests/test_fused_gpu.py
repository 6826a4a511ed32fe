"""LDS-resident 2-D steps: the multi-pass plan's fused shape, where one
workgroup per packet holds the packet's grid in LDS for a whole launch of
steps.  Its arithmetic is the row and strided pencil passes' (complex64:
pass_rows_pen_sp + pass_strided_pen; complex128: pass_rows_pen +
pass_strided_pen), so where both read the same V phase table psi and C(t)
must equal the multi-pass plan pinned to those kernels bit for bit, and
C(t) must match the FP64 oracle (|dC| <= 1e-6 mixed, 1e-10 FP64, as for
every multi-pass shape)."""
import numpy as np
import pytest

import espec_amd as E
from espec_amd import configs

pytestmark = pytest.mark.gpu


def packets(n, count, seed):
    """Packets on one n x n grid with their own masses, potentials and states."""
    rng = np.random.default_rng(seed)
    return [configs.packet2d(n, rng.uniform(0.1, 0.2, 2), rng.uniform(0.8, 1.4, 2),
                             rng.uniform(500, 2000, 2), rng.uniform(0.1, 0.5, 2),
                             rng.uniform(0.1, 0.3, 2), 0.5, 30) for _ in range(count)]


def plan_of(pk, copies, flags, lc, maxsamp):
    """A multi-pass plan holding `copies` repetitions of the packets pk."""
    n = pk[0].np[0] * pk[0].np[1]
    rep = lambda a: np.concatenate([a] * copies)
    return E.SpoPlan(pk[0].np, pk[0].dt_fs, lc, pk[0].sh,
                     rep(np.array([p.sm for p in pk], dtype=np.float64)),
                     rep(np.stack([p.vpot.reshape(n) for p in pk])),
                     rep(np.stack([p.u.reshape(n) for p in pk])),
                     rep(np.stack([p.v.reshape(n) for p in pk])),
                     maxsamp=maxsamp, flags=flags | E.EXACT | E.MULTIPASS)


def pencil_shape(flags, np0):
    """The multi-pass kernels whose arithmetic the fused shape repeats."""
    return dict(rows_pen=3 if flags == E.MIXED else 1, rows=1024 // np0, rows_nt=64,
                tile_pen=1, pen_pf=0, grid=0, fused=0)


def dC(g, o):
    k = min(len(g["cr"]), len(o["cr"]))
    return max(np.abs(g["cr"][:k] - o["cr"][:k]).max(), np.abs(g["ci"][:k] - o["ci"][:k]).max())


CASES = [(128, E.MIXED, 1e-6), (64, E.MIXED, 1e-6), (64, E.FP64, 1e-10)]


@pytest.mark.parametrize("nw", [4, 8])
@pytest.mark.parametrize("n,flags,tol", CASES)
def test_fused_matches_pencil_passes(gpu, orc, n, flags, tol, nw):
    """Four packets, C(t) every third step, a run of 7 steps then one of 13
    (two launches, the state in HBM between them)."""
    from oracle import EXACT
    pk = packets(n, 4, 31 + n)
    nst, lc = 20, 3
    print(f"case: fused {n}^2 flags={flags} nw={nw}", flush=True)
    ref = plan_of(pk, 1, flags, lc, nst // lc + 1)
    ref.set_shape(**pencil_shape(flags, n))
    fz = plan_of(pk, 1, flags, lc, nst // lc + 1)
    fz.set_shape(**dict(pencil_shape(flags, n), fused=nw))
    assert fz.shape()["fused"] == nw
    for plan in (ref, fz):
        plan.run(7)
        plan.run(nst - 7)
    for k, p in enumerate(pk):
        a, b = ref.get(k, psi=True), fz.get(k, psi=True)
        for key in ("cr", "ci", "u", "v"):
            assert np.array_equal(a[key], b[key]), (k, key, np.abs(a[key] - b[key]).max())
        o = orc.pspofft(p.np, p.dt_fs, 0.0, (nst - 1) * p.dt_fs, lc, p.sh, p.sm, p.u, p.v,
                        p.vpot, EXACT)
        assert dC(b, o) <= tol, k
    ref.close()
    fz.close()


def test_fused_table_above_the_limit(gpu, orc):
    """260 complex64 128^2 packets: above the 64 MB phase-table limit the
    multi-pass kernels form exp(-i cst2 V) from V, while the fused form
    reads a table of its own; both against the oracle, and each other."""
    from oracle import EXACT
    pk = packets(128, 5, 77)
    copies, nst = 52, 12
    print(f"case: fused 128^2 x{5 * copies} own phase table", flush=True)
    ref = plan_of(pk, copies, E.MIXED, 1, nst + 1)
    ref.set_shape(**pencil_shape(E.MIXED, 128))
    fz = plan_of(pk, copies, E.MIXED, 1, nst + 1)
    fz.set_shape(**dict(pencil_shape(E.MIXED, 128), fused=8))
    assert ref.shape()["vph"] == 0 and fz.shape()["vph"] == 1
    for plan in (ref, fz):
        plan.run(nst)
    for k in (0, 3, 5 * copies - 1):
        a, b = ref.get(k, psi=False), fz.get(k, psi=False)
        assert dC(a, b) <= 1e-6, k
        p = pk[k % 5]
        o = orc.pspofft(p.np, p.dt_fs, 0.0, (nst - 1) * p.dt_fs, 1, p.sh, p.sm, p.u, p.v,
                        p.vpot, EXACT)
        assert dC(b, o) <= 1e-6, k
    ref.close()
    fz.close()


def test_fused_refused_where_it_does_not_fit(gpu):
    """128^2 in complex128 needs 272 KiB of LDS; 256^2 has no instance."""
    for n, flags in ((128, E.FP64), (256, E.MIXED)):
        pk = packets(n, 1, 5)
        plan = plan_of(pk, 1, flags, 1, 4)
        with pytest.raises(E.EspecError):
            plan.set_shape(**dict(pencil_shape(flags, n), fused=4))
        plan.close()
