"""Batches of uneven packets on one GPU.

Packets are grouped by grid (shape and spacing); each group is one
device-resident plan (SpoPlan, SodPlan or SilPlan) holding all of its
packets with their own potentials, initial states and masses.  run()
enqueues every group on its own stream, so the groups execute side by
side; sync() waits for all of them.  graph(nsteps) instead records the
next nsteps steps of every group as one HIP graph (one branch per group),
which a single launch replays.  This replaces the reference's job farm
(tools&scrips/parallel.x), which starts one espec.x per packet.

A 2-D or 3-D group whose packets fit one workgroup's LDS can run either
on the LDS-resident kernel (one workgroup per packet, every step in LDS)
or on the multi-pass kernels (one pass per axis per step through HBM,
many workgroups per packet); with kind="auto" both plans are built and
timed over a few steps, and the faster one is kept.
"""
import time
from collections import OrderedDict

import numpy as np

from . import MIXED, EXACT, MULTIPASS, SpoBatchGraph, SpoPlan


def _step_time(plan, nsteps=40):
    """Seconds per step of `plan` (best of two timed runs), then reset."""
    plan.run(4)
    plan.sync()
    best = float("inf")
    for _ in range(2):
        t0 = time.perf_counter()
        plan.run(nsteps)
        plan.sync()
        best = min(best, (time.perf_counter() - t0) / nsteps)
    plan.reset()
    return best


class SpoBatch:
    """Split-operator propagation of a list of configs.Packet objects."""

    def __init__(self, packets, flags=MIXED | EXACT, maxsamp=None, lc=1, kind="auto"):
        """kind: "auto" (time the LDS-resident and multi-pass plans where
        both apply), "lds" (LDS-resident where the grid fits) or
        "multipass"."""
        if kind not in ("auto", "lds", "multipass"):
            raise ValueError(f"kind must be auto, lds or multipass, not {kind!r}")
        self.packets = packets
        self.lc = lc
        groups = OrderedDict()
        for i, p in enumerate(packets):
            key = (tuple(p.np), tuple(np.round(p.sh, 15)), p.dt_fs)
            groups.setdefault(key, []).append(i)
        # Plans (and their streams) are created smallest first, the largest
        # last: measured on the 2048-packet workload, a first 20-step run
        # took 0.31 ms per step in this order against 0.32-0.33 ms in order
        # of appearance and 0.36 ms largest first.
        order = sorted(groups, key=lambda k: (len(groups[k]) * int(np.prod(k[0])), k))
        self.groups = []
        self.where = {}
        for key in order:
            idx = groups[key]
            p0 = packets[idx[0]]
            n = int(np.prod(p0.np))
            ms = maxsamp if maxsamp is not None else p0.nsteps // lc + 1
            vp = np.stack([packets[i].vpot for i in idx]).reshape(len(idx), n)
            u = np.stack([packets[i].u for i in idx]).reshape(len(idx), n)
            v = np.stack([packets[i].v for i in idx]).reshape(len(idx), n)
            sm = np.array([packets[i].sm for i in idx], dtype=np.float64)

            def make(f):
                return SpoPlan(p0.np, p0.dt_fs, lc, p0.sh, sm, vp, u, v, maxsamp=ms, flags=f)

            plan = make(flags | (MULTIPASS if kind == "multipass" else 0))
            if kind == "auto" and len(p0.np) > 1 and plan.shape()["lds"]:
                alt = make(flags | MULTIPASS)
                if _step_time(alt) < 0.97 * _step_time(plan):
                    plan.close()
                    plan = alt
                else:
                    alt.close()
            self.groups.append(dict(np=list(p0.np), n=n, idx=idx, plan=plan))
            for k, i in enumerate(idx):
                self.where[i] = (len(self.groups) - 1, k)

    @property
    def points(self):
        """Grid points over all packets."""
        return sum(g["n"] * len(g["idx"]) for g in self.groups)

    def run(self, nsteps):
        for g in self.groups:
            g["plan"].run(nsteps)

    def sync(self):
        for g in self.groups:
            g["plan"].sync()

    def reset(self):
        for g in self.groups:
            g["plan"].reset()

    def timed_run(self, nsteps):
        """Wall time of nsteps steps of every packet (inputs resident)."""
        self.sync()
        t0 = time.perf_counter()
        self.run(nsteps)
        self.sync()
        return time.perf_counter() - t0

    def graph(self, nsteps):
        """Every group's next nsteps steps recorded as one graph
        (SpoBatchGraph): one launch then advances the whole batch."""
        return SpoBatchGraph([g["plan"] for g in self.groups], nsteps)

    def timed_graph_run(self, graph):
        """Wall time of one replay of `graph` (from graph()), inputs resident."""
        self.sync()
        t0 = time.perf_counter()
        graph.run()
        graph.sync()
        return time.perf_counter() - t0

    def get(self, i, psi=False):
        gi, k = self.where[i]
        return self.groups[gi]["plan"].get(k, psi=psi)

    def close(self):
        for g in self.groups:
            g["plan"].close()
        self.groups = []
