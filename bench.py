#!/usr/bin/env python3
"""Throughput of the split-operator FFT propagation on one MI355X.

Workload: the 256x256 coupled-Morse surface (two Morse modes, De=0.17 Eh,
a=1.0/bohr, masses 1836 and 918 m_e, bilinear coupling 0.005 q1 q2), product
Gaussian packet (sigma 0.2 bohr, displaced +0.3 bohr), dt = 1.0 a.u., C(t)
every step.  A "step" is one split-operator time step of the packet.

Reported: grid-point-steps/s of the mixed path (complex64 state in HBM,
FP64 arithmetic in the kernels) as `value`, the FP64 path beside it, the
largest |C(t) - C_oracle(t)| over the timed run, a roofline object for the
dominant kernel (HIP events on the plan's stream) and the FP64 C oracle
timed on the host as the CPU baseline.  Inputs are resident in HBM before
the timed region starts; only the propagation loop is timed.

Also reported, under "batched": the 2048-packet uneven batch (1-D and 2-D
grids, per-packet potentials, states and masses; 1000 steps of dt = 0.5
a.u.) on the mixed and FP64 paths, with its whole-job rate against the HBM
peak at 28 B per grid-point-step and a C(t) check of sampled packets.
Under "config3": the 128^3 surface, 1000 steps, mixed and FP64.
Under "krylov_controller": the learned Krylov-order/step controller
(trained on the device) against fixed-order short-iterative Lanczos on the
256x256 surface, wall time and |dC| against the order-20 run.

    python bench.py [--gpus 1] [--steps 2000] [--warmup 100] [--batch-steps 1000]
                    [--config3-steps 1000] [--krylov-steps 2000]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = "grid-point-steps/s, mixed vs FP64 SO-FFT on 1 GPU at |dC(t)|<=1e-6"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Launch shape of the 256x256 plans: one 256-point line per wavefront in
# both passes (radix-4 Stockham in registers), XCD-contiguous columns, one
# workgroup per line.  Pinned so that the kernels profiled by rocprofv3 are
# the ones timed here: the autotuner's pick between near-equal shapes can
# change with the load on the device (a profiler is load).
C2_SHAPE = dict(rows_pen=2, rows=1, rows_nt=64, tile_pen=3, pen_pf=0, xmap=1, grid=0, spread=1)


def timed_run(E, c, flags, steps, warmup):
    """Create a plan (inputs go to HBM), warm up, restart from psi0, time
    `steps` steps between synchronisations.  Returns (seconds, C(t), plan)."""
    plan = E.SpoPlan(c.np, c.dt_fs, 1, c.sh, c.sm, c.vpot, c.u, None, maxsamp=steps + 1,
                     flags=flags)
    plan.set_shape(**C2_SHAPE)
    if warmup:
        plan.run(warmup)
        plan.sync()
        plan.reset()
    plan.sync()
    t0 = time.perf_counter()
    plan.run(steps)
    plan.sync()
    t1 = time.perf_counter()
    res = plan.get(0, psi=False)
    return t1 - t0, res, plan


def pmc_traffic(name, kernel):
    """HBM bytes per dispatch of `kernel` (template instantiation name) from
    the committed rocprofv3 --pmc summary profiles/pmc_<name>.json, if it
    holds that instantiation (the autotuner may pick another shape)."""
    path = os.path.join(ROOT, "profiles", f"pmc_{name}.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            k = json.load(f)["kernels"].get(kernel)
        return None if k is None else k["traffic_bytes"]
    except (OSError, ValueError, KeyError):
        return None


def kernel_names(shape, np0, np1, storage="float"):
    """Instantiation names (as rocprofv3 prints them, namespace and
    arguments stripped) of the row-pass and strided-pass kernels that a
    launch shape selects for steps 2..K."""
    pht = "true" if shape.get("vph") else "false"
    if shape["rows_pen"] == 2:
        rows = f"pass_rows_wide<{storage}, {np0}, true, true, {pht}>"
    elif shape["rows_pen"]:
        rows = f"pass_rows_pen<{storage}, {np0 // 16}, true, true, {pht}>"
    else:
        ept_r = shape["rows"] * np0 // shape["rows_nt"]
        rows = f"pass_rows<{storage}, {shape['rows_nt']}, true, true, {ept_r}, {pht}>"
    if shape["tile_pen"] == 3:
        cols = f"pass_strided_wide<{storage}, {np1}>"
    elif shape["tile_pen"] == 2:
        cols = f"pass_strided_tile<{storage}, {np1 // 16}>"
    elif shape["tile_pen"]:
        pf = "true" if shape["pen_pf"] else "false"
        cols = f"pass_strided_pen<{storage}, {np1 // 16}, {pf}>"
    else:
        ept_s = shape["tile"] * np1 // shape["tile_nt"]
        cols = f"pass_strided<{storage}, {shape['tile_nt']}, {ept_s}>"
    return rows, cols


def batched(E, configs, steps, check):
    """The batched uneven-packet workload: 2048 seeded packets (1-D N in
    {128..1024}, 2-D 64^2..256^2), each with its own potential, state and
    masses, dt = 0.5 a.u., `steps` steps, C(t) every step.  Packets are
    grouped by grid into device-resident plans on their own streams; the
    timed region is one run of every group (inputs resident).  With
    `check`, the C(t) of one packet per 2-D class and of one 1-D packet is
    compared with the FP64 oracle over the whole run."""
    from espec_amd.batch import SpoBatch
    pk = configs.config4(count=2048, nsteps=steps, seed=4)
    out = {"packets": len(pk), "steps": steps, "lc": 1}
    for name, flags in (("mixed", E.MIXED | E.EXACT), ("fp64", E.FP64 | E.EXACT)):
        b = SpoBatch(pk, flags=flags)
        b.run(8)
        b.sync()
        b.reset()
        sec = b.timed_run(steps)
        gps = b.points * steps / sec
        out[name] = {"seconds": sec, "value": gps, "unit": "grid-point-steps/s",
                     "algorithmic_GBps": gps * 28.0 / 1e9 if name == "mixed" else gps * 56.0 / 1e9,
                     "groups": [{"grid": g["np"], "packets": len(g["idx"]),
                                 "shape": g["plan"].shape()} for g in b.groups]}
        log(f"[bench] batched {name}: {sec:.3f} s, {gps:.3e} gp-steps/s")
        if check and name == "mixed":
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            from oracle import Oracle, EXACT
            O = Oracle()
            picks = {}
            for i, p in enumerate(pk):
                key = tuple(p.np) if len(p.np) == 2 else ("1-D",)
                picks.setdefault(key, i)
            worst = {}
            for key, i in picks.items():
                p = pk[i]
                o = O.pspofft(p.np, p.dt_fs, 0.0, (steps - 1) * p.dt_fs, 1, p.sh, p.sm, p.u, p.v,
                              p.vpot, EXACT)
                g = b.get(i)
                n = min(len(o["cr"]), len(g["cr"]))
                d = max(np.abs(o["cr"][:n] - g["cr"][:n]).max(),
                        np.abs(o["ci"][:n] - g["ci"][:n]).max())
                worst["x".join(map(str, p.np))] = float(d)
            out["max_abs_dC_mixed_sample"] = worst
            log(f"[bench] batched mixed |dC| sample: {worst}")
        b.close()
    m = out["mixed"]["value"]
    out["step_roofline"] = {"bytes_per_gp_step": 28.0, "achieved_GBps": m * 28.0 / 1e9,
                            "frac": m * 28.0 / HBM_PEAK}
    out["mixed_over_fp64"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    return out


def config3(E, configs, steps, check):
    """The 128^3 three-mode surface (config 3 of the workload set): split-
    operator FFT, dt = 1 a.u., C(t) every step, `steps` steps on the mixed
    and FP64 paths, with inputs resident.  |dC| of the mixed run against
    the FP64 run over all steps and, with `check`, of both against the FP64
    oracle over the first 10 steps."""
    c = configs.config3(nsteps=steps)
    out = {"workload": "128^3 three-mode Morse split-operator FFT (3-D)", "steps": steps}
    res = {}
    for name, flags in (("mixed", E.MIXED | E.EXACT), ("fp64", E.FP64 | E.EXACT)):
        plan = E.SpoPlan(c.np, c.dt_fs, 1, c.sh, c.sm, c.vpot, c.u, None, maxsamp=steps + 1,
                         flags=flags)
        plan.run(40)
        plan.sync()
        plan.reset()
        t0 = time.perf_counter()
        plan.run(steps)
        plan.sync()
        sec = time.perf_counter() - t0
        res[name] = plan.get(0)
        gps = c.n * steps / sec
        bpp = 28.0 if name == "mixed" else 56.0
        out[name] = {"seconds": sec, "value": gps, "unit": "grid-point-steps/s",
                     "algorithmic_GBps": gps * bpp / 1e9, "frac_of_hbm_peak": gps * bpp / HBM_PEAK,
                     "shape": plan.shape()}
        log(f"[bench] config3 {name}: {sec:.3f} s, {gps:.3e} gp-steps/s")
        plan.close()
    out["max_abs_dC_mixed_vs_fp64"] = float(max(np.abs(res["mixed"]["cr"] - res["fp64"]["cr"]).max(),
                                                np.abs(res["mixed"]["ci"] - res["fp64"]["ci"]).max()))
    out["mixed_over_fp64"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    if check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle, EXACT
        o = Oracle().pspofft(c.np, c.dt_fs, 0.0, 9 * c.dt_fs, 1, c.sh, c.sm, c.u, c.v, c.vpot,
                             EXACT)
        k = len(o["cr"])
        out["max_abs_dC_vs_oracle_10_steps"] = {
            n: float(max(np.abs(res[n]["cr"][:k] - o["cr"]).max(),
                         np.abs(res[n]["ci"][:k] - o["ci"]).max())) for n in res}
    return out


def krylov_controller(E, configs, steps):
    """The learned Krylov-order/step controller on short-iterative Lanczos
    propagation of the 256x256 surface (FD Hamiltonian, dt = 1 a.u., C(t)
    every step, FP64).  The network is trained on the device from traces
    of 32 surfaces varied within +-20% around it (states advanced at orders
    12, 14, 16 and 21, 400 steps each), then `steps` adaptive steps
    (budget 1e-10 per dt, fallback order 12) are timed against the fixed
    orders 10 to 14; |dC| is measured against the order-20 run.  The
    budget per dt is 1e-6 / steps, so that the run stays within 1e-6 even
    if every step spends all of its budget."""
    from espec_amd import krylov
    tol = 1e-6 / steps
    t0 = time.perf_counter()
    fam = krylov.family(count=32, n=256, nsteps=0, seed=5, groups=4)
    traces = [krylov.trace_packets([p for p in fam if p.meta["group"] == g], 400, mprop=mp)
              for g, mp in zip(range(4), (12, 14, 16, 21))]
    x, t = krylov.dataset(traces, tol)
    params, hist = krylov.train(x, t, epochs=40, lr=3e-3, seed=3)
    t_train = time.perf_counter() - t0
    log(f"[bench] controller: {len(x)} samples, traces + training {t_train:.1f} s, "
        f"loss {hist[0]:.2e} -> {hist[-1]:.2e}")
    c = configs.config2(nsteps=steps)
    shm = krylov.shm_of(c)

    def plan(m):
        return E.SilPlan(c.np, m, 0.0, c.dt_fs, 1, shm, c.vpot, c.u, c.v, maxsamp=steps + 1)

    ref = plan(20)
    ref.run(steps)
    g20 = ref.get(0)
    ref.close()

    def dC(g):
        return float(max(np.abs(g["cr"] - g20["cr"]).max(), np.abs(g["ci"] - g20["ci"]).max()))

    out = {"workload": "256x256 coupled-Morse SIL, FD Hamiltonian, dt 1 a.u., FP64",
           "steps": steps, "budget_per_dt": tol, "train_samples": int(len(x)),
           "train_seconds": t_train, "fixed": {}}
    for m in (10, 11, 12, 13, 14):
        f = plan(m)
        f.run(16)  # warm-up: captures the step graph
        f.reset()
        f.sync()
        t1 = time.perf_counter()
        f.run(steps)
        f.sync()
        out["fixed"][str(m)] = {"seconds": time.perf_counter() - t1, "max_abs_dC": dC(f.get(0))}
        f.close()
    a = plan(21)
    a.adaptive(steps, params, tol=tol, m_fix=12)  # warm-up: captures the step graphs it meets
    a.reset()
    a.sync()
    t1 = time.perf_counter()
    st = a.adaptive(steps, params, tol=tol, m_fix=12)
    ta = time.perf_counter() - t1
    out["adaptive"] = {"seconds": ta, "max_abs_dC": dC(a.get(0)), "stats": st}
    a.close()
    ok = [m for m, v in out["fixed"].items() if v["max_abs_dC"] <= 1e-6]
    base = min(ok, key=int) if ok else None
    out["cheapest_fixed_within_1e-6"] = base
    out["speedup_vs_that_fixed"] = (out["fixed"][base]["seconds"] / ta) if base else None
    da = out["adaptive"]["max_abs_dC"]
    same = [m for m, v in out["fixed"].items() if v["max_abs_dC"] <= da]
    eq = min(same, key=int) if same else None
    out["cheapest_fixed_as_accurate"] = eq
    out["speedup_vs_that_fixed_as_accurate"] = (out["fixed"][eq]["seconds"] / ta) if eq else None
    log(f"[bench] controller: adaptive {ta:.2f} s |dC| {out['adaptive']['max_abs_dC']:.1e}; "
        f"fixed " + ", ".join(f"m={m} {v['seconds']:.2f} s |dC| {v['max_abs_dC']:.1e}"
                              for m, v in out["fixed"].items()))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and oracle check")
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--batch-steps", type=int, default=1000,
                    help="steps of the batched 2048-packet workload (0 skips it)")
    ap.add_argument("--krylov-steps", type=int, default=2000,
                    help="steps of the learned Krylov controller comparison (0 skips it)")
    ap.add_argument("--config3-steps", type=int, default=1000,
                    help="steps of the 128^3 workload (0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the propagation of one packet (or a batch of "
              "independent packets) runs on a single MI355X and no multi-GPU path is "
              "built; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import espec_amd as E
    from espec_amd import configs

    c = configs.config2(nsteps=args.steps)
    npts = c.n
    K, W = args.steps, args.warmup

    log(f"[bench] mixed path: {c.meta['name']} {K} steps (warmup {W})")
    sec_mix, res_mix, plan_mix = timed_run(E, c, E.MIXED | E.EXACT, K, W)
    ev_ms_mix, launches = plan_mix.timing()
    log(f"[bench] mixed: {sec_mix * 1e3:.2f} ms host, {ev_ms_mix:.2f} ms events, {launches} launches")
    log(f"[bench] FP64 path")
    sec_64, res_64, plan_64 = timed_run(E, c, E.FP64 | E.EXACT, K, W)
    log(f"[bench] fp64: {sec_64 * 1e3:.2f} ms host")

    # Per-kernel durations (timing-only HIP events on the plan's stream
    # around every launch of a 200-step run, no graph).
    plan_mix.reset()
    prof = plan_mix.profile(args.profile_steps)
    plan_mix.reset()
    shape = plan_mix.shape()
    inst = kernel_names(shape, c.np[0], c.np[1])
    descr = ["x pass: inverse x, V phases, C(t), forward x, Kx",
             "y pass: forward y, Ky, inverse y"]
    kinds = [dict(kernel=inst[i] if len(prof) > 1 else "spo_lds_fast",
                  role=descr[i] if len(prof) > 1 else "whole step in LDS", ms=m, bytes=b)
             for i, (m, b) in enumerate(prof)]
    # The row pass carries every byte of the per-grid-point-step model
    # (state read + write, V, psi0/LC: 2*8 + 4 + 8/LC for complex64); the
    # strided pass only adds the multi-pass round trip, which the model
    # does not count.  The roofline kernel is therefore the row pass (or
    # the single LDS-resident step kernel).
    dom = kinds[0]
    achieved = dom["bytes"] / (dom["ms"] * 1e-3)
    log(f"[bench] kernels: " + ", ".join(f"{k['kernel']} {k['ms'] * 1e3:.2f} us" for k in kinds))

    ms_bw, by_bw = E.bw_probe(1 << 30, 10)
    copy_bw = by_bw / (ms_bw * 1e-3)
    log(f"[bench] copy kernel: {copy_bw / 1e9:.0f} GB/s")

    out = {
        "metric": METRIC,
        "value": npts * K / sec_mix,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": sec_mix * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "storage": "complex64",
        "data": "synthetic",
        "config": {"workload": "256x256 coupled-Morse split-operator FFT (2-D), mixed path",
                   "grid": c.np, "dt_au": 1.0, "lc": 1, "steps": K,
                   "launches_per_run": launches, "launch_shape": plan_mix.shape(),
                   "launch_shape_pinned": True},
        "fp64": {"value": npts * K / sec_64, "ms_per_step": sec_64 * 1e3 / K,
                 "mixed_over_fp64": sec_64 / sec_mix},
        "kernels": kinds,
        "copy_kernel_GBps": copy_bw / 1e9,
        "roofline": {
            "bound": "hbm",
            "kernel": dom["kernel"],
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": pmc_traffic("config2_mixed", dom["kernel"]),
        },
        "timed_region_us_per_launch": ev_ms_mix * 1e3 / max(launches, 1),
        # whole-step view: gp-steps/s x (2*8 + 4 + 8/LC) bytes against the peak
        "step_roofline": {"bytes_per_gp_step": 28.0,
                          "achieved_GBps": npts * K / sec_mix * 28.0 / 1e9,
                          "frac": npts * K / sec_mix * 28.0 / HBM_PEAK},
    }

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle, EXACT
        O = Oracle()
        tf = (K - 1) * c.dt_fs  # the DO loop runs (TF-TI)/DT + 1 steps
        log(f"[bench] CPU oracle (FP64 C, 1 thread): {K} steps of the same packet")
        t0 = time.perf_counter()
        o = O.pspofft(c.np, c.dt_fs, 0.0, tf, 1, c.sh, c.sm, c.u, c.v, c.vpot, EXACT)
        t_cpu = time.perf_counter() - t0
        n = min(len(o["cr"]), len(res_mix["cr"]))
        dmix = max(np.abs(o["cr"][:n] - res_mix["cr"][:n]).max(),
                   np.abs(o["ci"][:n] - res_mix["ci"][:n]).max())
        d64 = max(np.abs(o["cr"][:n] - res_64["cr"][:n]).max(),
                  np.abs(o["ci"][:n] - res_64["ci"][:n]).max())
        out["max_abs_dC"] = {"mixed": float(dmix), "fp64": float(d64), "samples": int(n)}
        out["norm_drift"] = {"mixed": float(np.abs(res_mix["norm"][:n] - 1).max()),
                             "fp64": float(np.abs(res_64["norm"][:n] - 1).max())}
        out["cpu_baseline"] = {"value": npts * K / t_cpu, "unit": "grid-point-steps/s",
                               "cores": 1, "kind": "port",
                               "sample": f"the same 256x256 packet, {K} steps, FP64 C oracle "
                                         f"(oracle/espec_oracle.c), {t_cpu:.1f} s"}
        log(f"[bench] CPU: {t_cpu:.1f} s; max|dC| mixed {dmix:.2e} fp64 {d64:.2e}")
    plan_mix.close()
    plan_64.close()
    if args.batch_steps > 0:
        out["batched"] = batched(E, configs, args.batch_steps, not args.no_cpu)
    if args.config3_steps > 0:
        out["config3"] = config3(E, configs, args.config3_steps, not args.no_cpu)
    if args.krylov_steps > 0:
        out["krylov_controller"] = krylov_controller(E, configs, args.krylov_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
