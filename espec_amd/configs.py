"""Synthetic workloads: grids, potentials and initial packets.

Every quantity is in atomic units except time steps handed to the
propagators, which the eSPec interfaces take in fs (dt_fs = dt_au * SHFS,
pspofft.f:76-78).  Conventions that the workload descriptions leave open are
frozen here once, and tests and bench.py use only these functions:

* Gaussian width: psi0 = exp(-(x-x0)^2 / (2 sigma^2)), discretely normalised
  so that sum |psi|^2 = 1 (the reference normalises on the grid the same
  way, init_cond.f:60-65).
* No potential shift (SHIFTP off): V is used as given.
* 2-D/3-D grids span 8 bohr per axis, q in [-2, 6] bohr measured from each
  Morse minimum; 1-D batched packets likewise.
* Spectrum window: exp(-(t/tau)^2) (the .SG form of spectrumtd.f:65-76).
* Batched mix: half 1-D, half 2-D packets; sizes drawn uniformly from the
  listed sets with numpy's PCG64 and the stated seed.
"""
from dataclasses import dataclass, field

import numpy as np

SHFS = 2.4188843243e-2


@dataclass
class Packet:
    np: list            # grid points per axis
    sh: list            # grid step per axis (bohr)
    sm: list            # mass per axis (m_e)
    vpot: np.ndarray    # flattened potential, x fastest (hartree)
    u: np.ndarray       # Re psi0
    v: np.ndarray       # Im psi0
    dt_fs: float        # time step in fs
    nsteps: int
    lc: int = 1
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return int(np.prod(self.np))


def morse(q, de, a):
    return de * (1.0 - np.exp(-a * q)) ** 2


def gaussian(x, x0, sigma):
    return np.exp(-((x - x0) ** 2) / (2.0 * sigma ** 2))


def _normalise(psi):
    return psi / np.sqrt(np.sum(np.abs(psi) ** 2))


def config1(nsteps=4096):
    """1-D Morse, N=512 on [0.5, 6.0] bohr, m=918, psi0 at 1.8 bohr, sigma 0.15,
    dt = 0.5 a.u., C(t) every step."""
    n = 512
    x = np.linspace(0.5, 6.0, n)
    h = x[1] - x[0]
    v = morse(x - 1.4, 0.1744, 1.1)
    psi = _normalise(gaussian(x, 1.8, 0.15))
    return Packet([n], [h], [918.0], v, psi.real.copy(), np.zeros(n), 0.5 * SHFS, nsteps, 1,
                  dict(name="config1-morse1d-512"))


def _grid(n, lo=-2.0, hi=6.0):
    q = np.linspace(lo, hi, n)
    return q, q[1] - q[0]


def config2(nsteps=2000, n=256):
    """2-D coupled Morse, 256x256, De=0.17, a=1.0, masses 1836/918, coupling
    0.005 q1 q2, product Gaussian sigma 0.2 displaced +0.3, dt = 1.0 a.u."""
    q, h = _grid(n)
    q1, q2 = np.meshgrid(q, q, indexing="xy")  # q1 varies along x (fastest)
    v = morse(q1, 0.17, 1.0) + morse(q2, 0.17, 1.0) + 0.005 * q1 * q2
    psi = _normalise(gaussian(q1, 0.3, 0.2) * gaussian(q2, 0.3, 0.2))
    return Packet([n, n], [h, h], [1836.0, 918.0], v.ravel(), psi.ravel().copy(),
                  np.zeros(n * n), 1.0 * SHFS, nsteps, 1, dict(name=f"config2-morse2d-{n}x{n}"))


def config3(nsteps=1000, n=128, seed=3):
    """3-D: three Morse stretches (De U[0.15,0.2], a U[0.9,1.2]), pairwise
    bilinear couplings U[-0.01,0.01], masses U[1000,2000], Gaussian sigma
    0.25 displaced +0.3 per axis, dt = 1.0 a.u."""
    rng = np.random.Generator(np.random.PCG64(seed))
    de = rng.uniform(0.15, 0.2, 3)
    a = rng.uniform(0.9, 1.2, 3)
    c = rng.uniform(-0.01, 0.01, 3)
    m = rng.uniform(1000.0, 2000.0, 3)
    q, h = _grid(n)
    q1, q2, q3 = np.meshgrid(q, q, q, indexing="ij")
    # flatten with x fastest: index = i + n*j + n*n*k  ->  arrays [k][j][i]
    q1, q3 = q3, q1
    v = (morse(q1, de[0], a[0]) + morse(q2, de[1], a[1]) + morse(q3, de[2], a[2])
         + c[0] * q1 * q2 + c[1] * q1 * q3 + c[2] * q2 * q3)
    psi = _normalise(gaussian(q1, 0.3, 0.25) * gaussian(q2, 0.3, 0.25) * gaussian(q3, 0.3, 0.25))
    return Packet([n, n, n], [h, h, h], list(m), v.ravel(), psi.ravel().copy(),
                  np.zeros(n ** 3), 1.0 * SHFS, nsteps, 1, dict(name=f"config3-morse3d-{n}^3"))


def packet1d(n, de, a, m, x0, sigma, dt_au, nsteps, lc=1):
    q, h = _grid(n)
    v = morse(q, de, a)
    psi = _normalise(gaussian(q, x0, sigma))
    return Packet([n], [h], [m], v, psi.copy(), np.zeros(n), dt_au * SHFS, nsteps, lc)


def packet2d(n, de, a, m, x0, sigma, dt_au, nsteps, lc=1, coupling=0.0):
    q, h = _grid(n)
    q1, q2 = np.meshgrid(q, q, indexing="xy")
    v = morse(q1, de[0], a[0]) + morse(q2, de[1], a[1]) + coupling * q1 * q2
    psi = _normalise(gaussian(q1, x0[0], sigma[0]) * gaussian(q2, x0[1], sigma[1]))
    return Packet([n, n], [h, h], list(m), v.ravel(), psi.ravel().copy(), np.zeros(n * n),
                  dt_au * SHFS, nsteps, lc)


def config4(count=2048, nsteps=1000, seed=4):
    """Batched uneven packets: 1-D N in {128,256,512,1024}, 2-D in
    {64^2,128^2,256^2}; De U[0.1,0.2], a U[0.8,1.4], m U[500,2000], centre
    U[0.1,0.5] bohr from the minimum, sigma U[0.1,0.3]; dt = 0.5 a.u."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(count):
        if k % 2 == 0:
            n = int(rng.choice([128, 256, 512, 1024]))
            p = packet1d(n, rng.uniform(0.1, 0.2), rng.uniform(0.8, 1.4), rng.uniform(500, 2000),
                         rng.uniform(0.1, 0.5), rng.uniform(0.1, 0.3), 0.5, nsteps)
        else:
            n = int(rng.choice([64, 128, 256]))
            p = packet2d(n, rng.uniform(0.1, 0.2, 2), rng.uniform(0.8, 1.4, 2),
                         rng.uniform(500, 2000, 2), rng.uniform(0.1, 0.5, 2),
                         rng.uniform(0.1, 0.3, 2), 0.5, nsteps)
        p.meta["index"] = k
        out.append(p)
    return out
