"""Learned Krylov-order / time-step controller for the short-iterative
Lanczos propagator (no reference counterpart; the propagator itself is
PLNZ/PLNZAUX, plnz.f and sil.f).

Pieces:

* traces: Lanczos runs of 21 iterations per step on the device record,
  per step and packet, the controller inputs and, for each network order
  k = 4..20 and t = dt/2, dt, 2 dt, the estimate an adaptive step of order
  k + 1 is judged by: the larger of the residual estimate
  beta0 beta_{k+1} |[exp(-i t T_{k+1})]_{k+1,1}| and beta0 |y_k - y_{k+1}|,
  y_j = exp(-i t T_j) e1 (``SilPlan.trace``);
* labels: for each step, the cheapest (k, s) whose estimate meets the
  error budget (``targets``);
* training: the 8-64-64-2 bfloat16 network learns inputs of step k ->
  label of step k+1, with Adam on the device (``espec_amd.mlp_train``);
* use: ``SilPlan.adaptive(ndt, params)`` lets the network choose m and
  s = dt scale before every step; a step whose residual estimate misses
  the budget is redone with the fixed-order step (the fallback and judge);
* check: ``split_operator_check`` compares a finished run's C(t) with the
  split-operator FFT propagator's on the same packet.

Decision rule (the device host loop, sil.hip): k = ceil(4 + 16 o1) and
s = 2^round(2 o2 - 1), so the training targets are t1 = (m - 4) / 16 and
t2 = (log2 s + 1) / 2 for the labelled (m, s).  The step runs k + 2
iterations and keeps the smallest of the orders k-1..k+2 that meets the
budget (the inputs are the previous step's, and a miss costs a fallback).
"""
import numpy as np

from . import FP64, EXACT, SHFS, SilPlan, SpoPlan, MLP_NPAR, mlp_train, configs

M0, M1 = 4, 20
SCALES = (0.5, 1.0, 2.0)


def shm_of(p):
    """FD kinetic prefactor per axis, 1/(24 m h^2) (espec.f:630 with SHF = 1)."""
    return [1.0 / (24.0 * m * h * h) for m, h in zip(p.sm, p.sh)]


def targets(lab, tol, overhead=4.0):
    """Cheapest feasible (k, s) per step and packet from a trace's estimate
    table lab[nsteps][2 npk][17][3], as network targets [nsteps][npk][2].

    Budget per base step tol: s = 2 needs est(k, 2 dt) <= 2 tol; s = 1
    est(k, dt) <= tol; s = 1/2 runs two half steps, each est(k, dt/2) <=
    tol/2.  Cost per base step in Lanczos-iteration units: (k + 1 +
    overhead) per (sub)step, plus one for the second coefficient set of
    s = 2.  With nothing feasible the target is (20, 1/2)."""
    ns, nv = lab.shape[:2]
    err = lab.reshape(ns, nv // 2, 2, M1 - M0 + 1, 3).max(axis=2)  # worst of Re/Im
    ms = np.arange(M0, M1 + 1, dtype=np.float64)
    it = ms + 1.0  # iterations of a step of network order k
    best_cost = np.full(err.shape[:2], np.inf)
    tm = np.full(err.shape[:2], float(M1))
    ts = np.full(err.shape[:2], 0.5)
    for s in SCALES:
        if s == 0.5:
            ok = err[..., 0] <= tol / 2
            cost = 2.0 * (it + overhead)
        elif s == 1.0:
            ok = err[..., 1] <= tol
            cost = it + overhead
        else:
            ok = err[..., 2] <= 2 * tol
            cost = (it + overhead + 1.0) / 2.0
        c = np.where(ok, cost, np.inf)            # [ns][npk][17]
        k = c.argmin(axis=-1)
        cmin = np.take_along_axis(c, k[..., None], axis=-1)[..., 0]
        better = cmin < best_cost
        best_cost = np.where(better, cmin, best_cost)
        tm = np.where(better, ms[k], tm)
        ts = np.where(better, s, ts)
    return np.stack([(tm - M0) / 16.0, (np.log2(ts) + 1.0) / 2.0], axis=-1)


def family(count=32, n=128, nsteps=3000, seed=5, groups=4):
    """Training surfaces: config-2-style coupled 2-D Morse surfaces with De,
    a, coupling, the packet's offset and width varied within +-20% of
    config 2's values, dt = 1 a.u.  The masses are varied per group of
    count/groups packets (a Lanczos plan shares its kinetic prefactor)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u = lambda k=None: rng.uniform(0.8, 1.2, k)
    out = []
    for g in range(groups):
        m = [1836.0 * u(), 918.0 * u()]
        for _ in range(count // groups):
            p = configs.packet2d(n, 0.17 * u(2), 1.0 * u(2), m, 0.3 * u(2), 0.2 * u(2), 1.0,
                                 nsteps, coupling=0.005 * u())
            p.meta["group"] = g
            out.append(p)
    return out


def family1d(count=32, nsteps=0, seed=7, groups=4):
    """Training surfaces for the 1-D Morse workload: its grid (512 points on
    [0.5, 6.0] bohr), De, a, the minimum, the packet's offset from it and
    its width varied within +-20% of 0.1744 Eh, 1.1 /bohr, 1.4 bohr, 0.4
    bohr and 0.15 bohr, dt = 0.5 a.u.; the mass (918 m_e +-20%) per group
    of count/groups packets."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u = lambda: rng.uniform(0.8, 1.2)
    x = np.linspace(0.5, 6.0, 512)
    h = x[1] - x[0]
    out = []
    for g in range(groups):
        m = 918.0 * u()
        for _ in range(count // groups):
            xe = 1.4 * u()
            v = configs.morse(x - xe, 0.1744 * u(), 1.1 * u())
            psi = configs.gaussian(x, xe + 0.4 * u(), 0.15 * u())
            psi = psi / np.sqrt(np.sum(psi * psi))
            p = configs.Packet([512], [h], [m], v, psi, np.zeros(512), 0.5 * SHFS, nsteps, 1)
            p.meta["group"] = g
            out.append(p)
    return out


def split_operator_reference(packet, nsamp, flags=FP64 | EXACT):
    """C(t) of `packet` from the split-operator FFT propagator, nsamp
    samples every dt (sample 0 = C(0))."""
    p = SpoPlan(packet.np, packet.dt_fs, 1, packet.sh, packet.sm, packet.vpot, packet.u,
                packet.v, maxsamp=nsamp, flags=flags)
    try:
        p.run(nsamp - 1)
        return p.get(0, psi=False)
    finally:
        p.close()


def split_operator_check(packet, res, flags=FP64 | EXACT, ref=None):
    """Independent check of a Lanczos propagation of `packet` (C(t) samples
    every dt in res["cr"], res["ci"], sample 0 = C(0)): the same packet
    propagated by the split-operator FFT propagator (spectral kinetic
    operator, PSPOFFT's step) over the same samples; returns
    max_t |C(t) - C_spo(t)|.  The Lanczos propagator applies the
    fourth-order finite-difference H of AU, so a converged run departs
    from the split-operator one by the two discretisations' difference
    only; a failed Krylov step shows as a departure well beyond that.
    ref: a split_operator_reference result to reuse."""
    ns = len(res["cr"])
    g = ref if ref is not None else split_operator_reference(packet, ns, flags)
    k = min(ns, len(g["cr"]))
    return float(max(np.abs(np.asarray(res["cr"][:k]) - g["cr"][:k]).max(),
                     np.abs(np.asarray(res["ci"][:k]) - g["ci"][:k]).max()))


def trace_packets(packets, nsteps, mp=21, mprop=None, flags=FP64 | EXACT):
    """Traces of packets sharing grid, masses and dt: (inputs, estimates);
    the state advances with order mprop (default mp)."""
    p0 = packets[0]
    n = p0.n
    plan = SilPlan(p0.np, mp, 0.0, p0.dt_fs, 1, shm_of(p0),
                   np.stack([p.vpot.reshape(n) for p in packets]),
                   np.stack([p.u.reshape(n) for p in packets]),
                   np.stack([p.v.reshape(n) for p in packets]), maxsamp=nsteps + 1, flags=flags)
    try:
        return plan.trace(nsteps, mprop)
    finally:
        plan.close()


def dataset(traces, tol, overhead=4.0):
    """Pairs (inputs of step k, targets of step k+1) from a list of traces."""
    xs, ts = [], []
    for feat, lab in traces:
        tg = targets(lab, tol, overhead)
        xs.append(feat[:-1].reshape(-1, 8))
        ts.append(tg[1:].reshape(-1, 2))
    return np.concatenate(xs).astype(np.float32), np.concatenate(ts).astype(np.float32)


def init_params(seed=0):
    """Glorot-uniform weights, zero biases."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = np.zeros(MLP_NPAR, dtype=np.float32)
    o = 0
    for fan_in, fan_out in ((8, 64), (64, 64), (64, 2)):
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        p[o:o + fan_in * fan_out] = rng.uniform(-lim, lim, fan_in * fan_out)
        o += fan_in * fan_out + fan_out
    return p


def train(x, t, epochs=40, lr=2e-3, seed=0):
    """Train from scratch on the device; returns (params, per-epoch loss)."""
    return mlp_train(init_params(seed), x, t, epochs, lr=lr, seed=seed + 1)
