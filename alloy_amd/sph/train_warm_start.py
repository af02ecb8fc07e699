"""Train the learned warm start of the pressure solves and write its weights.

    python -m alloy_amd.sph.train_warm_start [--samples 4000000] [--steps 30000] [--out alloy_amd/sph/data]

Data: (features -> correction to the warm-start rule) pairs that plain DFSPH logs while it steps
(``learned.collect``), from dam breaks with randomised block shapes (each side U{12..40} lattice sites, the tank
1.6 x the block plus 0.1 m, a seed per scene, 600 steps: collapse, run-up and sloshing) and from packs of batched
bowl worlds (16 worlds each, pack seeds 100+; the benchmark pack, seed 2, is never used). Held out for validation:
the 8,000- and 125,000-particle dam breaks and 32 worlds of the seed-2 bowl pack.
Training: Adam on minibatches with a cosine learning rate, every forward, backward, reduction and update a HIP
kernel of this library. Validation: the held-out scenes stepped plain and with the network attached, for each
choice of the solves it feeds (DFSPH_OPT_LEARNED_SOLVES, 0 = none); the choice with the fewest mean iterations over
the validation scenes is stored with the weights. Writes ``warm_start_mlp.npy`` (float32 parameters) and
``warm_start_mlp.json`` (the run's record).
"""

from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np
import torch

from . import _abi, scenes
from .learned import DEFAULT_WEIGHTS, WarmStartMLP, collect

DT = scenes.DT


def training_scenes(dam_breaks: int, bowl_packs: int, seed: int):
    """(scene, steps) pairs: randomised dam breaks, then 16-world bowl packs."""
    rng = np.random.default_rng(seed)
    for k in range(dam_breaks):
        nx, ny, nz = (int(v) for v in rng.integers(12, 41, 3))
        box = float(max(nx, ny, nz) * 0.01 * 1.6 + 0.1)
        yield scenes.dam_break(0, box, seed=1000 + seed * 100 + k, device="cuda", counts=(nx, ny, nz)), 600
    for k in range(bowl_packs):
        yield scenes.bowl_worlds(16, seed=100 + seed * 10 + k, device="cuda"), 400


def collect_dataset(samples: int, seed: int = 0, dam_breaks: int = 8, bowl_packs: int = 2, log=print):
    total_steps = 600 * dam_breaks + 400 * bowl_packs
    per_step = max(1, samples // total_steps)
    xs, ts = [], []
    for k, (sc, steps) in enumerate(training_scenes(dam_breaks, bowl_packs, seed)):
        s = sc.make_solver()
        st = sc.model.state()
        X, T = collect(s, st, steps=steps, dt=DT, per_step=per_step, seed=seed * 100 + k, warmup=5)
        xs.append(X)
        ts.append(T)
        log(f"  scene {k}: {sum(sc.fluid_counts)} particles, {sc.model.world_count} worlds -> {X.shape[0]} samples")
        del s, st
    return torch.cat(xs), torch.cat(ts)


def run_iterations(scene, steps: int, net: WarmStartMLP | None = None, solves: int = 3) -> dict:
    s = scene.make_solver()
    if net is not None:
        net.attach(s)
        s.set_option(_abi.OPT_LEARNED_SOLVES, solves)
    st = scene.model.state()
    it_d, it_v, den = [], [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(st, st, None, None, DT)
        r = s.report()
        it_d.append(float(np.mean(r.iters_density)))
        it_v.append(float(np.mean(r.iters_divergence)))
        den.append(float(np.mean(r.density_err)))
    torch.cuda.synchronize()
    return {"iters_density": float(np.mean(it_d)), "iters_divergence": float(np.mean(it_v)),
            "iters_total": float(np.mean(it_d) + np.mean(it_v)), "density_err_mean": float(np.mean(den)),
            "density_err_max": float(np.max(den)), "seconds": time.perf_counter() - t0}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--samples", type=int, default=4_000_000)
    ap.add_argument("--steps", type=int, default=30000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--eval-steps", type=int, default=400)
    ap.add_argument("--out", default=os.path.dirname(DEFAULT_WEIGHTS))
    args = ap.parse_args(argv)
    t0 = time.time()
    log = lambda m: print(f"[{time.time() - t0:7.1f}s] {m}", flush=True)  # noqa: E731
    log(f"collecting {args.samples} samples")
    X, T = collect_dataset(args.samples, args.seed, log=log)
    log(f"{X.shape[0]} samples; target mean {T.mean(0).tolist()} std {T.std(0).tolist()}")
    net = WarmStartMLP("cuda", seed=args.seed)
    losses = net.train(X, T, steps=args.steps, batch=args.batch, lr=args.lr, seed=args.seed,
                       log_every=max(1, args.steps // 10))
    base = float(((T - T.mean(0)) ** 2).mean().item()) / 2.0
    log(f"loss {losses[0]:.5f} -> {losses[-1]:.5f} (predicting the mean: {base:.5f})")
    evals = {}
    for name, sc in (("dam_break_8k", scenes.dam_break(20, 0.4, 0, device="cuda")),
                     ("dam_break_125k", scenes.dam_break(50, 1.0, 1, device="cuda")),
                     ("bowl_worlds_32_of_512", scenes.bowl_worlds(512, seed=2, device="cuda", subset=range(0, 512, 16)))):
        res = {"plain": run_iterations(sc, args.eval_steps)}
        for solves in (1, 2, 3):
            res[f"learned_{solves}"] = run_iterations(sc, args.eval_steps, net, solves)
        evals[name] = res
        log(f"{name}: " + ", ".join(f"{k} {v['iters_density']:.2f}/{v['iters_divergence']:.2f}"
                                     for k, v in res.items()))
    # relative iteration change per validation scene, averaged; 0 (no learned start) unless a choice helps
    score = {0: 0.0}
    for s in (1, 2, 3):
        score[s] = float(np.mean([evals[w][f"learned_{s}"]["iters_total"] / evals[w]["plain"]["iters_total"] - 1.0
                                  for w in evals]))
    best = min(score, key=score.get)
    os.makedirs(args.out, exist_ok=True)
    net.save(os.path.join(args.out, "warm_start_mlp.npy"))
    record = {"samples": int(X.shape[0]), "train_steps": args.steps, "batch": args.batch, "lr": args.lr,
              "seed": args.seed, "losses": losses, "mean_predictor_loss": base, "eval_steps": args.eval_steps,
              "eval": evals, "score_relative_iterations": score, "learned_solves": best,
              "data": "dam breaks with randomised block shapes (sides U{12..40}), 600 steps each, and 16-world bowl "
                      "packs (seeds 100+), 400 steps each: features and rule corrections of random liquid particles "
                      "per step"}
    with open(os.path.join(args.out, "warm_start_mlp.json"), "w") as f:
        json.dump(record, f, indent=1)
    log(f"saved to {args.out} (learned solves {best})")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
