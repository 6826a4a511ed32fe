"""Batches of uneven packets on one GPU.

Packets are grouped by grid (shape and spacing); each group is one
device-resident plan (SpoPlan, SodPlan or SilPlan) holding all of its
packets with their own potentials, initial states and masses.  run()
enqueues every group on its own stream, so the groups execute side by
side; sync() waits for all of them.  This replaces the reference's job
farm (tools&scrips/parallel.x), which starts one espec.x per packet.
"""
import time
from collections import OrderedDict

import numpy as np

from . import MIXED, EXACT, SpoPlan


class SpoBatch:
    """Split-operator propagation of a list of configs.Packet objects."""

    def __init__(self, packets, flags=MIXED | EXACT, maxsamp=None, lc=1):
        self.packets = packets
        self.lc = lc
        groups = OrderedDict()
        for i, p in enumerate(packets):
            key = (tuple(p.np), tuple(np.round(p.sh, 15)), p.dt_fs)
            groups.setdefault(key, []).append(i)
        self.groups = []
        self.where = {}
        for key, idx in groups.items():
            p0 = packets[idx[0]]
            n = int(np.prod(p0.np))
            ms = maxsamp if maxsamp is not None else p0.nsteps // lc + 1
            vp = np.stack([packets[i].vpot for i in idx]).reshape(len(idx), n)
            u = np.stack([packets[i].u for i in idx]).reshape(len(idx), n)
            v = np.stack([packets[i].v for i in idx]).reshape(len(idx), n)
            sm = np.array([packets[i].sm for i in idx], dtype=np.float64)
            plan = SpoPlan(p0.np, p0.dt_fs, lc, p0.sh, sm, vp, u, v, maxsamp=ms, flags=flags)
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

    def get(self, i, psi=False):
        gi, k = self.where[i]
        return self.groups[gi]["plan"].get(k, psi=psi)

    def close(self):
        for g in self.groups:
            g["plan"].close()
        self.groups = []
