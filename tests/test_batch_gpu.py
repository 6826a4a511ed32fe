"""Batched uneven packets (per-packet masses, potentials and initial states;
grids grouped into plans) against the oracle packet by packet."""
import numpy as np
import pytest

import espec_amd as E
from espec_amd import configs
from espec_amd.batch import SpoBatch

pytestmark = pytest.mark.gpu


def dC(g, o):
    k = min(len(g["cr"]), len(o["cr"]))
    return max(np.abs(g["cr"][:k] - o["cr"][:k]).max(), np.abs(g["ci"][:k] - o["ci"][:k]).max())


def family(n2d, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        if n2d:
            out.append(configs.packet2d(n2d, rng.uniform(0.1, 0.2, 2), rng.uniform(0.8, 1.4, 2),
                                        rng.uniform(500, 2000, 2), rng.uniform(0.1, 0.5, 2),
                                        rng.uniform(0.1, 0.3, 2), 0.5, 40))
        else:
            out.append(configs.packet1d(128, rng.uniform(0.1, 0.2), rng.uniform(0.8, 1.4),
                                        rng.uniform(500, 2000), rng.uniform(0.1, 0.5),
                                        rng.uniform(0.1, 0.3), 0.5, 40))
    return out


@pytest.mark.parametrize("n2d", [0, 64, 128])
@pytest.mark.parametrize("flags,tol", [(E.FP64 | E.EXACT, 1e-10), (E.MIXED | E.EXACT, 1e-6)])
def test_per_packet_masses(gpu, orc, n2d, flags, tol):
    from oracle import EXACT
    pk = family(n2d, 6, 11 + n2d)
    nst = 40
    print(f"case: batch n2d={n2d} flags={flags} packets={len(pk)}", flush=True)
    b = SpoBatch(pk, flags=flags)
    assert len(b.groups) == 1
    b.run(nst)
    b.sync()
    for i, p in enumerate(pk):
        o = orc.pspofft(p.np, p.dt_fs, 0.0, (nst - 1) * p.dt_fs, 1, p.sh, p.sm, p.u, p.v, p.vpot,
                        EXACT)
        assert dC(b.get(i), o) <= tol, i
    b.close()


@pytest.mark.parametrize("kind", ["lds", "multipass", "auto"])
@pytest.mark.parametrize("n2d", [0, 64])
@pytest.mark.parametrize("flags,tol", [(E.FP64 | E.EXACT, 1e-10), (E.MIXED | E.EXACT, 1e-6)])
def test_plan_kinds(gpu, orc, kind, n2d, flags, tol):
    """Packets that fit one workgroup's LDS on the LDS-resident kernel, on
    the multi-pass kernels (one-dimensional grids: the step counter is
    advanced without a strided pass) and on the timed choice of the two;
    C(t) every second step."""
    from oracle import EXACT
    pk = family(n2d, 5, 31 + n2d)
    nst, lc = 40, 2
    print(f"case: plan kinds kind={kind} n2d={n2d} flags={flags}", flush=True)
    b = SpoBatch(pk, flags=flags, lc=lc, kind=kind)
    if kind != "auto":
        assert b.groups[0]["plan"].shape()["lds"] == (kind == "lds")
    b.run(nst)
    b.sync()
    for i, p in enumerate(pk):
        o = orc.pspofft(p.np, p.dt_fs, 0.0, (nst - 1) * p.dt_fs, lc, p.sh, p.sm, p.u, p.v,
                        p.vpot, EXACT)
        g = b.get(i)
        assert len(g["cr"]) == nst // lc + 1
        assert dC(g, o) <= tol, i
    b.close()


def test_config4_subset(gpu, orc):
    """The uneven mix of the batched configuration (1-D N in 128..1024, 2-D
    64^2..256^2), a 48-packet sample, mixed path."""
    from oracle import EXACT
    pk = configs.config4(count=48, nsteps=60, seed=4)
    print(f"case: config4 subset {len(pk)} packets", flush=True)
    b = SpoBatch(pk, flags=E.MIXED | E.EXACT)
    b.run(60)
    b.sync()
    for i in range(0, len(pk), 5):
        p = pk[i]
        o = orc.pspofft(p.np, p.dt_fs, 0.0, 59 * p.dt_fs, 1, p.sh, p.sm, p.u, p.v, p.vpot, EXACT)
        assert dC(b.get(i), o) <= 1e-6, (i, p.np)
    b.close()


@pytest.mark.parametrize("flags", [E.MIXED | E.EXACT, E.FP64 | E.EXACT])
def test_batch_graph_matches_streams(gpu, flags):
    """The uneven mix (LDS-resident 1-D groups, multi-pass 2-D groups) run
    as one graph launch gives the per-plan stream runs' C(t) and state bit
    for bit; replaying needs the plans back at the captured step."""
    pk = configs.config4(count=40, nsteps=37, seed=4)
    nst = 37
    print(f"case: batch graph vs streams, {len(pk)} packets, flags={flags}", flush=True)
    b = SpoBatch(pk, flags=flags, maxsamp=nst + 1)
    b.run(nst)
    b.sync()
    ref = [b.get(i, psi=True) for i in range(len(pk))]
    b.reset()
    g = b.graph(nst)
    for rep in range(2):
        g.run()
        ms = g.sync()
        assert ms > 0
        for i in range(len(pk)):
            r = b.get(i, psi=True)
            for k in ("cr", "ci", "norm", "u", "v"):
                assert np.array_equal(r[k], ref[i][k]), (rep, i, k)
        with pytest.raises(E.EspecError):
            g.run()  # the plans have moved past the captured steps
        b.reset()
    g.close()
    b.close()
