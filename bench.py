#!/usr/bin/env python3
"""Benchmark of the REMC hot path on one MI355X.

Workload: the 40-chain crowded box (lengths ~ U[20,300], about 6.4k residues,
400 uM total molecule concentration, box 549.6 A), 32 replicas at 288-450 K,
rigid translate/rotate moves (2 A / 0.2 rad), exchange every 1000 MC steps.
A "step" is one MC step of every replica: on the device, each replica picks a
molecule, proposes a move, evaluates the delta-E row of that molecule against
all others (FP32 pair terms, Kahan, FP64 reduction) and applies Metropolis.
Exactly --warmup steps are run untimed and exactly --steps steps are timed:
whole exchange intervals (every --exchange-every steps an exchange round)
are replayed from a captured hipGraph, a remainder is launched directly.

value = MC steps/s summed over all replicas (replicas x steps / wall), with
the replicas already resident in HBM.  Before the warm-up the autotuner
picks the pair kernel's layout and launch shape on the pending proposals.
The roofline object refers to the dominant kernel (mc_row_kernel in the
j-lane layout, mc_pair_kernel in the i-lane layout): algorithmic FP32 flops
= 35 per inter-molecular residue pair, pairs counted on the device, divided
by the kernel's mean duration measured with HIP events recorded by the
dispatches themselves.  Its HBM
traffic per launch comes from two rocprofv3 --pmc passes (FETCH_SIZE,
WRITE_SIZE), scaled by what the same counters read for a plain device copy
of known size (--copy-probe-mib), and is read from the committed summary in
profiles/ when present.

The delayed_acceptance object compares plain REMC with the learned
surrogate's delayed acceptance on the same workload from the same state
(effective samples/s, stage-one acceptance; see da_leg).

The cpu_baseline leg times the reference's CPU loop (full Replica::E per MC
step, pthreads over replicas) as restated in oracle/, on a bounded sample,
with one thread per core of the process's CPU share (at most 16, the GPU
box's share for one GPU) and with one thread.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FLOPS_PER_PAIR = 35            # algorithmic FP32 flops per residue pair
PEAK_FP32_TFLOPS = 157.3       # MI355X dense FP32 vector peak (packed FMA)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50000, help="timed MC steps (all replicas each)")
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--config", type=int, default=1, choices=[0, 1, 2, 4],
                    help="0: two chains, 1: 40-chain crowded box (default), 2: linked domains (flexible), "
                         "4: 256-chain crowded box")
    ap.add_argument("--exchange-every", type=int, default=1000)
    ap.add_argument("--replicas", type=int, default=0, help="override the configuration's replica count")
    ap.add_argument("--wg-per-cu", type=int, default=0, help="workgroups per CU of the pair kernel (0: tuned)")
    ap.add_argument("--layout", default="auto", choices=["auto", "ilanes", "jlanes", "jlanes4", "resident"],
                    help="pair-kernel layout (auto: tuned)")
    ap.add_argument("--no-autotune", action="store_true", help="skip the pair-kernel autotuner")
    ap.add_argument("--window", type=int, default=0, choices=[0, 1, 2, 4, 8],
                    help="MC steps per speculative window (0: automatic, 1: one step per launch pair)")
    ap.add_argument("--no-image-shift", action="store_true", help="per-pair minimum image everywhere")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline sample length")
    ap.add_argument("--no-da", action="store_true", help="leave out the delayed-acceptance leg")
    ap.add_argument("--da-equilibrate", type=int, default=400000,
                    help="plain MC steps before the delayed-acceptance comparison (training samples are "
                         "collected along them)")
    ap.add_argument("--da-steps", type=int, default=1200000, help="MC steps of each leg of the comparison")
    ap.add_argument("--da-train-steps", type=int, default=3000)
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="only run this many direct MC launches (for rocprofv3), no JSON line")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this size (counter calibration)")
    return ap.parse_args()


KERNEL_NAMES = {"ilanes": "mc_pair_kernel", "jlanes": "mc_row_kernel", "jlanes4": "mc_row_kernel",
                "resident": "mc_resident_kernel"}


def committed_traffic(kernel, workload, key="bytes_per_launch"):
    """HBM bytes per pair-kernel launch (key bytes_per_launch_w<window
    slots>) or, for the resident kernel, whose one launch runs many MC
    steps, per MC step (key bytes_per_step), from the committed PMC summary
    profiles/<kernel>_hbm_traffic_<workload>.txt (line
    '<key> <workload> <value>'), or None."""
    path = ROOT / "profiles" / f"{kernel}_hbm_traffic_{workload}.txt"
    try:
        for line in path.read_text().splitlines():
            f = line.split()
            if len(f) == 3 and f[0] == key and f[1] == workload:
                return float(f[2])
    except OSError:
        pass
    return None


def copy_probe(mib):
    """A plain copy of known size: FETCH_SIZE / WRITE_SIZE of this kernel
    calibrate the counters (each should read mib MiB)."""
    import torch

    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x 3", flush=True)
    return 0


# CPU pair rate per thread of the oracle's reference loop (order of magnitude,
# used only to decide whether the CPU legs fit the sample budget)
CPU_PAIRS_PER_S_PER_THREAD = 4.0e7


def cpu_cores():
    """Cores this process may run on (its affinity mask), at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(spec, seed, target_s):
    """The reference's CPU loop on a bounded sample (oracle restatement): one
    thread per core of the process's share, then one thread."""
    import oracle

    threads = cpu_cores()
    st = oracle.Remc(spec, seed=seed)
    # calibrate with one MC step per replica, then scale to the target
    t1 = st.reference_run(1, threads)
    n = max(1, int(target_s / max(t1, 1e-6)))
    t = st.reference_run(n, threads)
    steps = n * spec.n_rep
    # one thread, on one replica for a quarter of the sample length (the
    # loop runs a thread's replicas in turn, so the rate is per replica step)
    one = oracle.Remc(spec, n_rep=1, seed=seed)
    t1s = one.reference_run(1, 1)
    n1 = max(1, int(target_s / 4 / max(t1s, 1e-6)))
    t_one = one.reference_run(n1, 1)
    return {
        "value": steps / t,
        "unit": "MC steps/s (all replicas)",
        "cores": threads,
        "kind": "port",
        "cpu_model": cpu_model(),
        "sample": f"{n} MC steps x {spec.n_rep} replicas of the same workload, full Replica::E per step, "
                  f"{threads} pthreads, {t:.1f} s",
        "pairs_per_s": steps * spec.pairs_full() / t,
        "one_thread": {"value": n1 / t_one, "unit": "MC steps/s (all replicas)",
                       "sample": f"{n1} MC steps of one replica of the same workload, 1 thread, {t_one:.1f} s"},
    }


def tau_int(x, c=5.0):
    """Integrated autocorrelation time, Sokal's automatic window (c = 5)."""
    from cgppd_amd.surrogate_study import tau_int as t
    return t(x, c)


def da_leg(args, layout, wg_per_cu):
    """Plain and delayed-acceptance REMC on the 40-chain box from the same
    state: after --da-equilibrate plain steps (training pairs collected
    along them every 250 steps) and --da-train-steps Adam steps, both legs
    run --da-steps steps from that state, recording every replica's energy
    after each 1000-step exchange interval.  Effective samples/s = sum over
    the temperatures of n / tau_int of the energy at that temperature,
    divided by the leg's wall time."""
    from cgppd_amd import engine, synth

    spec = synth.config(1)
    per = 1000
    e = engine.Engine(0)
    e.load_spec(spec, seed=args.seed + 7)
    e.set_pair_layout(layout)
    e.set_launch(wg_per_cu)
    e.mc_prepare()
    e.surrogate_init(seed=args.seed)
    t0 = time.perf_counter()
    for _ in range(max(1, args.da_equilibrate // per)):
        for _ in range(4):
            e.mc_steps(per // 4)
            e.surrogate_collect()
        e.exchange()
    n_train = e.surrogate_sample_count()
    t_eq = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = [e.surrogate_train(args.da_train_steps // 10, batch=1024, lr=1e-3 if k < 7 else 5e-4, seed=args.seed + k)
              for k in range(10)]
    e.synchronize()
    t_train = time.perf_counter() - t0
    params = e.surrogate_params()
    temps = e.stats()["temperature"]
    labels = e.stats()["label"]
    cen = np.zeros((spec.n_rep, spec.n_mol, 3), np.float32)
    rel = np.zeros((spec.n_rep, spec.n_res, 3), np.float32)
    for r in range(spec.n_rep):
        _, rel[r], cen[r] = e.coordinates(r)
    e.close()
    print(f"delayed acceptance: {n_train} training pairs over {t_eq:.1f} s, loss {losses[0]:.3g} -> {losses[-1]:.3g} "
          f"in {t_train:.1f} s", file=sys.stderr, flush=True)

    legs = {}
    for name, da in (("plain", False), ("delayed_acceptance", True)):
        f = engine.Engine(0)
        f.load_spec(spec, seed=args.seed + 7)
        ts, rs = spec.step_sizes()
        f.set_replicas(cen, rel, temps, ts, rs, labels, seed=args.seed + 11)
        f.set_pair_layout(layout)
        f.set_launch(wg_per_cu)
        f.mc_prepare()
        if da:
            f.surrogate_set_params(params)
            f.set_delayed_acceptance(True)
        f.run(0, per, use_graph=True)
        s0 = f.stats()
        rej0 = f.stage1_rejections().sum()
        n = max(1, args.da_steps // per)
        energies = []
        f.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            f.run(1, per, use_graph=True)
            s = f.stats()
            energies.append(s["energy"][np.argsort(s["temperature"], kind="stable")])
        wall = time.perf_counter() - t0
        energies = np.array(energies)
        taus = np.array([tau_int(energies[:, k]) for k in range(spec.n_rep)])
        ess = float((n / taus).sum())
        acc = s["counters"].sum(0) - s0["counters"].sum(0)
        rej = int(f.stage1_rejections().sum() - rej0)
        decided = int(acc.sum()) + rej
        legs[name] = {
            "mc_steps_per_s": spec.n_rep * n * per / wall,
            "acceptance": float((acc[0] + acc[1]) / max(1, decided)),
            "stage1_acceptance": float(1.0 - rej / max(1, decided)) if da else None,
            "tau_int_intervals": {"min": float(taus.min()), "median": float(np.median(taus)),
                                  "max": float(taus.max())},
            "runs_cover_20_tau": bool(n >= 20 * taus.max()),
            "effective_samples": ess,
            "effective_samples_per_s": ess / wall,
            "wall_s": wall,
        }
        f.close()
        print(f"{name}: {legs[name]['mc_steps_per_s']:.0f} MC steps/s, ESS/s {legs[name]['effective_samples_per_s']:.3g}",
              file=sys.stderr, flush=True)
    p, d = legs["plain"], legs["delayed_acceptance"]
    return {
        "workload": spec.name,
        "steps_per_leg": (args.da_steps // per) * per,
        "equilibration_steps": max(1, args.da_equilibrate // per) * per,
        "training": {"pairs": int(n_train), "adam_steps": args.da_train_steps, "loss_first": losses[0],
                     "loss_last": losses[-1], "train_s": t_train},
        "plain": p,
        "delayed_acceptance": d,
        "ess_per_s_ratio": d["effective_samples_per_s"] / max(p["effective_samples_per_s"], 1e-30),
        "mc_steps_per_s_ratio": d["mc_steps_per_s"] / max(p["mc_steps_per_s"], 1e-30),
    }


def energy_error(e, spec):
    """Running totals vs a fresh FP64 CPU evaluation of replica 0."""
    s = e.stats()
    pos, _, _ = e.coordinates(0)
    want = fp64_energy(spec, pos)
    return abs(s["energy"][0] - want) / max(abs(want), 1e-30)


def fp64_energy(spec, pos):
    """FP64 oracle energy (kcal/mol) of one replica's coordinates: the
    inter-molecular pairs plus, for flexible molecules, their internal
    energy (intra-molecular pairs and bonded terms)."""
    import oracle

    args = (spec.mol_off, spec.types, spec.ff.charge, spec.ff.sigma, spec.ff.eps)
    want = oracle.total(*oracle.energy_fp64(*args, pos, spec.box))
    if getattr(spec, "flexible", False):
        fl = oracle.Flex(spec, spec.torsion_v, spec.torsion_delta)
        for m in range(spec.n_mol):
            if fl.mol_flexible[m]:
                want += oracle.internal_total(oracle.internal(fl, m, *args, pos, spec.box, 0, True))
    return want


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        return 2
    if args.copy_probe_mib:
        return copy_probe(args.copy_probe_mib)
    from cgppd_amd import engine, synth

    spec = synth.config(args.config)
    if args.replicas:
        spec.n_rep = args.replicas
    e = engine.Engine(0)
    e.load_spec(spec, seed=args.seed)
    if args.wg_per_cu:
        e.set_launch(args.wg_per_cu)
    if args.layout != "auto":
        e.set_pair_layout(args.layout)
    if args.no_image_shift:
        e.set_image_shift(False)
    e.set_window(args.window)
    e.mc_prepare()
    R = spec.n_rep
    tuned = None
    if not args.no_autotune and not (args.wg_per_cu and args.layout != "auto"):
        # the autotuner times the candidates on the pending proposals; the
        # caller's explicit choices stay in force
        tuned = e.autotune(launches=20)
        print(f"autotune: {json.dumps(tuned)}", file=sys.stderr, flush=True)
    layout = e.pair_layout()

    if args.profile_steps:
        e.mc_steps(args.profile_steps)
        e.synchronize()
        print(f"profiled {args.profile_steps} MC launches", flush=True)
        return 0

    per = args.exchange_every

    def advance(n):
        """n MC steps of every replica: whole exchange intervals from the
        captured graph, then the remainder as direct launches."""
        if n // per:
            e.run(n // per, per, use_graph=True)
        if n % per:
            e.mc_steps(n % per)

    # the interval graph is captured outside the timed region
    e.run(0, per, use_graph=True)
    advance(args.warmup)
    e.synchronize()

    steps = args.steps
    s0 = e.stats()
    p0 = e.pair_evaluations()
    e.synchronize()
    t0 = time.perf_counter()
    advance(steps)
    e.synchronize()
    wall = time.perf_counter() - t0
    pairs = e.pair_evaluations() - p0
    s1 = e.stats()

    # kernel timing pass: direct launches bracketed by HIP events on the
    # context stream (the pair kernel: each launch; the resident kernel, which
    # runs all the steps in one launch: its span per MC step)
    e.set_timing(True)
    p1 = e.pair_evaluations()
    n_timed = 2000 if layout == "resident" else 500
    e.mc_steps(n_timed)
    kern_ms, launches = e.last_kernel_ms()
    pairs_timed = e.pair_evaluations() - p1
    e.set_timing(False)
    # per launch of the row kernel (a window launch evaluates the rows of
    # several MC steps; the pairs are the algorithmic ones, n_m (N - n_m)
    # per step, without the energies between a window's proposals), or,
    # for the resident kernel, per MC step
    per_unit = n_timed if layout == "resident" else max(1, launches)
    achieved = pairs_timed / per_unit * FLOPS_PER_PAIR / (kern_ms * 1e-3) / 1e12

    # acceptance of the timed steps (counters follow the temperatures, so
    # the sums over replicas are unaffected by exchanges)
    acc = s1["counters"].sum(0) - s0["counters"].sum(0)
    ex = s1["exchange"] - s0["exchange"]
    line = {
        "metric": "MC steps/s (all replicas), plain delta-E REMC; 1e-5 rel energy parity; % FP32 VALU",
        "value": R * steps / wall,
        "unit": "MC steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": spec.name,
            "n_mol": spec.n_mol, "n_res": spec.n_res, "replicas": R, "box_A": round(spec.box, 2),
            "exchange_every": per,
            "moves": ("translate 2 A / rotate 0.2 rad / flex / local translate + crankshaft"
                      if getattr(spec, "flexible", False) else "rigid translate 2 A / rotate 0.2 rad"),
            "pair_terms": "fp32, plain sums of 16 terms into Kahan sums per lane, fp64 row reduction",
            "flexible": bool(getattr(spec, "flexible", False)),
            "image_shift": not args.no_image_shift,
            "pair_layout": layout,
            "window": e.window(),
            "wg_per_cu": e.pair_wg_per_cu(),
            "autotuned": tuned is not None,
        },
        "pairs_per_s": pairs / wall,
        "timed_region": {
            "acceptance": float((acc[0] + acc[1]) / max(1, acc.sum())),
            "exchange_rounds": int(steps // per),
            "exchange_acceptance": float(ex[1] / max(1, ex[0])),
            "state": (f"starts {args.warmup} MC steps after the synthetic dispersed placement (no "
                      f"equilibration run): the crowded boxes relax over ~1e5-1e6 steps; the work per "
                      f"step depends on the state only through the share of near visits (j-blocks within "
                      f"the largest LJ switch distance of a moved molecule: under 0.1 % of the 40-chain "
                      f"box's pair work at the start, about 6 % after 1e6 steps, where the rate measured "
                      f"0.6 % lower)"),
        },
        "roofline": {
            "bound": "valu",
            "kernel": KERNEL_NAMES[layout],
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": committed_traffic(KERNEL_NAMES[layout], spec.name,
                                         "bytes_per_step" if layout == "resident"
                                         else f"bytes_per_launch_w{e.window()}"),
            "kernel_ms": kern_ms,
            "pairs_per_launch": pairs_timed / per_unit,
            "mc_steps_per_launch": n_timed / per_unit,
            "flops_per_pair": FLOPS_PER_PAIR,
            # the resident kernel runs all n_timed steps in one launch: its
            # time, pairs and traffic are given per MC step
            "per": "MC step" if layout == "resident" else "launch",
        },
    }
    if not args.no_cpu_baseline:
        # the CPU legs evaluate every pair of a replica at least once: beyond
        # a few times the target sample length they are left out (config 4
        # would need ~10 minutes for its first calibration step)
        threads = cpu_cores()
        est_s = spec.pairs_full() * spec.n_rep / (threads * CPU_PAIRS_PER_S_PER_THREAD)
        if est_s <= 4.0 * args.cpu_seconds:
            line["energy_rel_err_vs_fp64"] = energy_error(e, spec)
            line["cpu_baseline"] = cpu_baseline(spec, args.seed, args.cpu_seconds)
        else:
            line["cpu_baseline"] = None
            line["cpu_baseline_skipped"] = (f"one full evaluation of every replica would take about {est_s:.0f} s "
                                            f"on {threads} threads")
    e.close()
    if not args.no_da and args.config == 1:
        line["delayed_acceptance"] = da_leg(args, layout, line["config"]["wg_per_cu"])
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
