"""Inputs of the reference's regression tests, rebuilt from their input.spc.

Only what the propagations read: the grid, the final potential and the
initial state.  Constants and formulas follow the reference:
  grid       XI + (I-1)*SH, SH = (XF-XI)/(NP-1) in angstrom, /A0A  (espec.f:345,619-622)
  mass       amu * 1836.152667539                                (espec.f:614)
  .OHS       V = k/2 (x-re)^2, k = (w c tau 2 pi)^2 m            (makecst.f:84-86, potfuncs.f:63-64)
  .MORSE     V = D (1-exp(-a(x-re)))^2, D = 2 pi c tau w^2/(4 wx),
             a = sqrt(k/(2D))                                    (makecst.f:141-148, potfuncs.f:104-105)
  SHM        1/(24 m h^2)                                        (espec.f:630)
  psi0       lowest eigenvector of the FD Hamiltonian of the initial
             potential; the reference builds the same matrix (ap1d.f)
             and diagonalises it (mtrxdiag.f) or runs LANCZS on it.
The regression outputs are eSPec v0.4 (results.test12:2), which predates
the potential shift by the grid minimum (SHIFTP, gvpot.f:64-73); the
potentials here are therefore not shifted.
"""
import gzip
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
A0A = 0.5291772083
AMU = 1836.152667539
CCGS = 2.99792458e10
TAU = 2.4188843243e-17
TWOPI = 6.2831853071795864769252867663


def load(name):
    with gzip.open(os.path.join(HERE, name + ".txt.gz"), "rt") as f:
        return np.array([[float(x.replace("D", "E")) for x in ln.split()] for ln in f if ln.strip()])


def grid(n, xi_a, xf_a):
    sh_a = (xf_a - xi_a) / (n - 1.0)
    xi, sh = xi_a / A0A, sh_a / A0A
    return xi + np.arange(n) * sh, sh


def ohs(x, re_a, w, m):
    k = (w * CCGS * TAU * TWOPI) ** 2 * m
    return 0.5 * k * (x - re_a / A0A) ** 2


def morse_cm(x, re_a, w, wx, m):
    k = (w * CCGS * TAU * TWOPI) ** 2 * m
    d = TWOPI * CCGS * TAU * w ** 2 / (4.0 * wx)
    a = np.sqrt(k / (2.0 * d))
    return d * (1.0 - np.exp(-a * (x - re_a / A0A))) ** 2


def fd_matrix(v, shm):
    n = len(v)
    h = np.diag(30.0 * shm + v) + np.diag(np.full(n - 1, -16.0 * shm), 1) \
        + np.diag(np.full(n - 2, shm), 2)
    return np.triu(h) + np.triu(h, 1).T


def ground_state(v, shm):
    w, z = np.linalg.eigh(fd_matrix(v, shm))
    psi = z[:, 0]
    return psi if psi[np.argmax(np.abs(psi))] > 0 else -psi


def test_case(t):
    """Return dict(np_, sh, sm, shm, vpot, u, v, dt, ti, tf, mf, method, mp)."""
    m = 7.49 * AMU
    if t == 14:
        # .PPSIL (PPLNZ) without a laser field: test12's surfaces on 201
        # points, LANCZS ground state of the initial surface, ABSTOL 1e-6.
        n = 201
        x, sh = grid(n, 0.5, 2.8)
        shm = 1.0 / (24.0 * m * sh * sh)
        v0 = ohs(x, 1.146, 1943.79, m)
        v1 = ohs(x, 1.211, 1629.23, m)
        psi = ground_state(v0, shm)
        return dict(np_=[n], sh=[sh], sm=[m], shm=[shm], vpot=v1, u=psi, v=np.zeros(n),
                    dt=2e-3, ti=0.0, tf=250.0, mf=12, method="PPSIL", mp=10, abstol=1e-6)
    if t in (12, 10):
        n = 256 if t == 12 else 250
        x, sh = grid(n, 0.5, 2.8)
        shm = 1.0 / (24.0 * m * sh * sh)
        v0 = ohs(x, 1.146, 1943.79, m)
        v1 = ohs(x, 1.211, 1629.23, m)
        psi = ground_state(v0, shm)
        method = "PFSPO" if t == 12 else "PFSOD"
        return dict(np_=[n], sh=[sh], sm=[m], shm=[shm], vpot=v1, u=psi, v=np.zeros(n),
                    dt=1e-4, ti=0.0, tf=250.0, mf=12, method=method, wp=1e-3)
    if t in (6, 5):
        n = 256
        x, sh = grid(n, 0.5, 2.8)
        shm = 1.0 / (24.0 * m * sh * sh)
        v0 = morse_cm(x, 1.146, 1943.79, 13.71, m)
        v1 = morse_cm(x, 1.211, 1629.23, 17.74, m)
        psi = ground_state(v0, shm)
        tf = 150.0 if t == 6 else 250.0
        return dict(np_=[n], sh=[sh], sm=[m], shm=[shm], vpot=v1, u=psi, v=np.zeros(n),
                    dt=5e-5, ti=0.0, tf=tf, mf=14, method="PFSPO" if t == 6 else "PFSOD", wp=1e-5)
    if t == 4:
        ps = load("eigpot_test3_psi")
        pv = load("eigpot_test3_pot")
        n = len(ps)
        # The .FILE grid is in bohr already ('R1_i/a.u. = 0.151178E+01',
        # results.test4): SH = (XF-XI)/(NP-1) without the angstrom factor.
        sh = (pv[-1, 0] - pv[0, 0]) / (n - 1.0)
        shm = 1.0 / (24.0 * m * sh * sh)
        psi = ps[:, 1].copy()  # used as read (.GETR); its norm prints as 1.0000002
        return dict(np_=[n], sh=[sh], sm=[m], shm=[shm], vpot=pv[:, 1].copy(), u=psi,
                    v=np.zeros(n), dt=2e-3, ti=0.0, tf=250.0, mf=12, method="PSIL", mp=10,
                    abstol=1e-6)
    raise KeyError(t)
