"""Surrogate delayed-acceptance study on one GPU.

    python -m cgppd_amd.surrogate_study [--config 1] [--samples 1000000] ...

1. Plain REMC runs on the device generate training data: every molecule
   pair of every replica with its exact pair energy, collected every
   --collect-every MC steps until --samples pairs are held on the device.
2. The surrogate is trained on the device (BF16 MFMA kernels, Adam) and
   checked on a held-out collection (RMSE, correlation).
3. Plain and delayed-acceptance REMC run the same procedure from the same
   start: --warmup steps, then --steps steps in chunks of --chunk steps,
   exchanges every --exchange-every steps; after each chunk the energy at
   every temperature rank and the inter-centre distances of the replica
   nearest 300 K are recorded.
4. Reported per run: MC steps/s, stage-one acceptance, the integrated
   autocorrelation time of the interaction energy at every temperature
   (Sokal's automatic window, c = 5), effective samples/s = sum over
   temperatures of n / tau_int divided by the wall time of the recorded
   steps, and the agreement of the interaction-energy and inter-centre
   distance histograms near 300 K (two-sample Kolmogorov-Smirnov).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def tau_int(x, c=5.0):
    """Integrated autocorrelation time with Sokal's automatic window:
    tau(W) = 1 + 2 sum_{t=1..W} rho(t), smallest W with W >= c tau(W)."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    if n < 4 or np.var(x) == 0:
        return 1.0
    y = x - x.mean()
    f = np.fft.rfft(y, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n]
    acf /= acf[0]
    tau = 1.0
    for w in range(1, n):
        tau += 2.0 * acf[w]
        if w >= c * tau:
            return max(tau, 1.0)
    return max(tau, 1.0)


def ks_2samp(a, b, tau_a=1.0, tau_b=1.0):
    """Two-sample Kolmogorov-Smirnov statistic D and its p-value for the
    effective sample sizes n / tau (correlated samples; the asymptotic
    Kolmogorov distribution of D sqrt(n_a n_b / (n_a + n_b)))."""
    from scipy import stats
    d = float(stats.ks_2samp(a, b).statistic)
    na, nb = len(a) / max(tau_a, 1.0), len(b) / max(tau_b, 1.0)
    p = float(stats.kstwobign.sf(d * np.sqrt(na * nb / (na + nb))))
    return {"D": d, "p_value": p, "n_eff": [na, nb]}


def record_run(e, spec, warmup, steps, chunk, exchange_every, da):
    """Runs warmup + steps MC steps (exchange every exchange_every) and
    records energies by temperature rank and inter-centre distances of the
    replica nearest 300 K after each chunk."""
    R, M = spec.n_rep, spec.n_mol
    temps = np.sort(e.stats()["temperature"])
    k300 = int(np.argmin(np.abs(temps - 300.0)))

    clock = {"steps": 0}

    def advance(n):
        done = 0
        while done < n:
            into = clock["steps"] % exchange_every
            k = min(n - done, exchange_every - into, chunk)
            e.mc_steps(k)
            done += k
            clock["steps"] += k
            if clock["steps"] % exchange_every == 0:
                e.exchange()

    advance(warmup)
    e.synchronize()
    energies, dists = [], []
    iu = np.triu_indices(M, 1)
    s0 = e.stats()
    rej0 = e.stage1_rejections().sum() if da else 0
    t0 = time.perf_counter()
    for _ in range(steps // chunk):
        advance(chunk)
        s = e.stats()
        order = np.argsort(s["temperature"], kind="stable")
        energies.append(s["energy"][order])
        _, _, cen = e.coordinates(int(order[k300]))
        d = cen[:, None, :] - cen[None, :, :]
        d -= spec.box * np.rint(d / spec.box)
        dists.append(np.linalg.norm(d, axis=2)[iu])
    e.synchronize()
    wall = time.perf_counter() - t0
    s1 = e.stats()
    energies = np.array(energies)
    taus = [tau_int(energies[:, k]) for k in range(R)]
    n = energies.shape[0]
    ess = float(sum(n / t for t in taus))
    acc = s1["counters"].sum(0) - s0["counters"].sum(0)
    rej1 = (e.stage1_rejections().sum() - rej0) if da else 0
    total = acc.sum() + rej1
    out = {
        "wall_s": wall,
        "mc_steps_per_s": R * (steps // chunk) * chunk / wall,
        "acceptance": float((acc[0] + acc[1]) / max(1, total)),
        "stage1_acceptance": float(1.0 - rej1 / max(1, total)) if da else None,
        "tau_int": [float(t) for t in taus],
        "samples_per_temperature": int(n),
        "effective_samples": ess,
        "effective_samples_per_s": ess / wall,
        "temperature_300K": float(temps[k300]),
    }
    return out, energies[:, k300], np.concatenate(dists)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--samples", type=float, default=1e6)
    ap.add_argument("--collect-every", type=int, default=25)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--warmup", type=int, default=200000)
    ap.add_argument("--steps", type=int, default=500000)
    ap.add_argument("--chunk", type=int, default=100)
    ap.add_argument("--exchange-every", type=int, default=1000)
    ap.add_argument("--translate-step", type=float, default=None, help="override the configuration's step")
    ap.add_argument("--rotate-step", type=float, default=None)
    args = ap.parse_args(argv)

    from . import engine, synth

    spec = synth.config(args.config)
    if args.translate_step is not None:
        spec.translate_step = args.translate_step
    if args.rotate_step is not None:
        spec.rotate_step = args.rotate_step
    e = engine.Engine(0)

    # 1. training data from plain REMC
    t0 = time.perf_counter()
    e.load_spec(spec, seed=args.seed + 1)
    e.mc_prepare()
    e.surrogate_init(seed=args.seed)
    e.mc_steps(args.warmup)
    while e.surrogate_sample_count() < args.samples:
        e.run(1, args.collect_every, use_graph=True)
        e.surrogate_collect()
    n_train = e.surrogate_sample_count()
    t_collect = time.perf_counter() - t0
    print(f"collected {n_train} training pairs in {t_collect:.1f} s", file=sys.stderr, flush=True)

    # 2. training and a held-out check
    t0 = time.perf_counter()
    losses = []
    for k in range(10):
        lr = args.lr * (0.5 if k >= 7 else 1.0)
        losses.append(e.surrogate_train(args.train_steps // 10, batch=args.batch, lr=lr, seed=args.seed + k))
    e.synchronize()
    t_train = time.perf_counter() - t0
    print(f"trained {args.train_steps} steps in {t_train:.1f} s, loss {losses[0]:.3g} -> {losses[-1]:.3g}",
          file=sys.stderr, flush=True)
    e.run(4, args.collect_every, use_graph=True)
    # held out: the pairs within the surrogate's range (centre distance below
    # r_I + r_J + 15 A, bounding radii of the relative positions)
    held_x, held_t = [], []
    iu = np.triu_indices(spec.n_mol, 1)
    radius = np.array([np.linalg.norm(spec.rel[spec.mol_off[m]:spec.mol_off[m + 1]], axis=1).max()
                       for m in range(spec.n_mol)])
    for r in range(spec.n_rep):
        _, _, cen = e.coordinates(r)
        d = cen[:, None, :] - cen[None, :, :]
        d -= spec.box * np.rint(d / spec.box)
        inr = (np.linalg.norm(d, axis=2) < radius[:, None] + radius[None, :] + 15.0)[iu]
        held_x.append(e.surrogate_features(r)[inr])
        held_t.append(e.energy_matrix(r)[iu][inr])
    held_x = np.concatenate(held_x)
    held_t = np.clip(np.concatenate(held_t), -15, 15)
    pred = e.surrogate_predict(held_x)
    near = np.abs(held_t) > 0.05
    params = e.surrogate_params()

    result = {
        "workload": spec.name, "replicas": spec.n_rep, "n_mol": spec.n_mol, "n_res": spec.n_res,
        "translate_step": spec.translate_step, "rotate_step": spec.rotate_step,
        "training": {
            "samples": int(n_train), "collect_s": t_collect, "train_steps": args.train_steps, "batch": args.batch,
            "train_s": t_train, "loss_first": losses[0], "loss_last": losses[-1],
            "heldout_pairs_in_range": int(len(held_t)),
            "heldout_rmse_kcal": float(np.sqrt(np.mean((pred - held_t) ** 2))),
            "heldout_rmse_interacting_kcal": float(np.sqrt(np.mean((pred[near] - held_t[near]) ** 2)))
            if near.any() else None,
            "heldout_corr_interacting": float(np.corrcoef(pred[near], held_t[near])[0, 1]) if near.sum() > 2 else None,
        },
    }
    e.close()

    # 3. plain and delayed-acceptance runs from the same start
    runs = {}
    for name, da in (("plain", False), ("delayed_acceptance", True)):
        e = engine.Engine(0)
        e.load_spec(spec, seed=args.seed + 2)
        e.mc_prepare()
        if da:
            e.surrogate_set_params(params)
            e.set_delayed_acceptance(True)
        runs[name] = record_run(e, spec, args.warmup, args.steps, args.chunk, args.exchange_every, da)
        e.close()
        print(f"{name}: {runs[name][0]['mc_steps_per_s']:.0f} MC steps/s", file=sys.stderr, flush=True)
    plain, pe, pd = runs["plain"]
    dar, de, dd = runs["delayed_acceptance"]
    result["plain"] = plain
    result["delayed_acceptance"] = dar
    result["ess_per_s_ratio"] = dar["effective_samples_per_s"] / max(plain["effective_samples_per_s"], 1e-30)
    result["steps_per_s_ratio"] = dar["mc_steps_per_s"] / max(plain["mc_steps_per_s"], 1e-30)
    M = spec.n_mol
    npair = M * (M - 1) // 2
    result["histogram_energy_300K_ks"] = ks_2samp(pe, de, tau_int(pe), tau_int(de))
    # distances: npair per recorded sample, correlated through the energy's tau
    result["histogram_com_distance_300K_ks"] = ks_2samp(pd, dd, npair * tau_int(pe), npair * tau_int(de))
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
