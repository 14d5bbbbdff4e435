"""Learned warm start: collect training data from the plain solver, train
the MLP on the GPU, evaluate it against the zero start.

    python -m shapes_amd.warmstart collect --samples 5000000 --out data.npz
    python -m shapes_amd.warmstart train --data data.npz --out warmstart_mlp.npy
    python -m shapes_amd.warmstart eval --model warmstart_mlp.npy

Training data: (features, targets) of every contact that missed the cache,
logged after the plain solver (zero start, coloured FP32, 8 sweeps) on
polygon-pile worlds (dt 1/60; the committed warmstart_mlp.npy) or on
batches of piles (`--scene pile_batch`, dt 1/120, logged after `--skip`
settling steps) whose seeds differ from the evaluation seeds.  Targets
are the solver's final impulses times the rows' inverse effective masses
(velocity units).  The model is trained with Adam on mini-batches by the
library's own BF16 matrix-core forward/backward kernels.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import engine, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "warmstart_mlp.npy")
EVAL_SEED = 1002        # the benchmark's polygon pile
TRAIN_SEEDS_FROM = 5001


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal hidden weights, small output layer, zero biases."""
    rng = np.random.default_rng(seed)
    shapes = [(14, 128), (128,), (128, 128), (128,), (128, 128), (128,), (128, 2), (2,)]
    parts = []
    for k, s in enumerate(shapes):
        if len(s) == 2:
            w = rng.normal(0.0, np.sqrt(2.0 / s[0]), s)
            if k == 6:
                w *= 0.1
            parts.append(w.ravel())
        else:
            parts.append(np.zeros(s))
    return np.concatenate(parts).astype(np.float32)


def collect(samples: int, count: int = 8192, steps: int = 600, iterations: int = 8,
            solver: int = engine.SOLVER_COLORED_F32, scene: str = "polygon_pile",
            worlds: int = 64, skip: int = 0, log=print):
    feats, tgts = [], []
    total = 0
    seed = TRAIN_SEEDS_FROM
    while total < samples:
        if scene == "pile_batch":
            sc = scenes.pile_batch(seed=seed, n_worlds=worlds, iterations=iterations)
            b = engine.make_batch(sc.worlds, sc.config, solver=solver,
                                  broadphase=engine.BROADPHASE_GRID_1, max_pairs_per_body=4)
        else:
            sc = scenes.polygon_pile(seed=seed, count=count, steps=steps, iterations=iterations)
            b = engine.make_batch(sc.worlds, sc.config, solver=solver)
        if skip:
            b.step_async(skip, use_graph=True)
            b.sync()
        b.set_logging(samples - total)
        t0 = time.perf_counter()
        b.step(steps)
        f, t = b.read_log()
        b.close()
        feats.append(f)
        tgts.append(t)
        total += len(f)
        log(f"collect: seed {seed}: {len(f)} samples in {time.perf_counter() - t0:.1f} s "
            f"(total {total})")
        seed += 1
    f = np.concatenate(feats)[:samples]
    t = np.concatenate(tgts)[:samples]
    # the on-device log order depends on atomics; sort for a reproducible set
    order = np.lexsort(np.concatenate([f, t], axis=1).T[::-1])
    return f[order], t[order]


def train(f, t, epochs: int = 8, batch: int = 4096, lr: float = 1e-3, seed: int = 0,
          huber: float = 0.0, clip: float = 20.0, log=print):
    """Adam on mini-batches; targets clipped to +-clip (0: no clip); loss
    squared error, or Huber with threshold `huber`."""
    if clip > 0:
        t = np.clip(t, -clip, clip)
    net = engine.Mlp(init_params(seed))
    net.set_loss(huber)
    net.set_data(f, t)
    rng = np.random.default_rng(seed)
    n = (len(f) // batch) * batch
    hist = []
    for ep in range(epochs):
        idx = rng.permutation(len(f))[:n]
        lr_ep = lr * (0.5 ** max(0, ep - epochs // 2))
        t0 = time.perf_counter()
        loss = net.train(idx, batch, lr_ep, want_loss=(ep == epochs - 1 or ep == 0))
        dt = time.perf_counter() - t0
        if loss is not None:
            hist.append(float(loss.mean()))
            log(f"train: epoch {ep} lr {lr_ep:.2e} mean batch loss {loss.mean():.5f} ({dt:.2f} s)")
        else:
            log(f"train: epoch {ep} lr {lr_ep:.2e} ({dt:.2f} s)")
    p = net.params()
    net.close()
    return p, hist


def evaluate(params, seed: int = EVAL_SEED, iters=(2, 4, 8), steps: int = 600,
             count: int = 8192, log=print):
    """Zero start vs learned start on the polygon pile with the coloured
    solver: run maxima and run means of the per-step maxima of penetration
    and residuals, and body-steps/s."""
    out = []
    for it in iters:
        for mode in ("zero", "learned"):
            sc = scenes.polygon_pile(seed=seed, count=count, steps=steps, iterations=it)
            ws = engine.WARM_LEARNED if mode == "learned" else engine.WARM_ZERO
            b = engine.make_batch(sc.worlds, sc.config, solver=engine.SOLVER_COLORED_F32,
                                  warm_start=ws)
            if mode == "learned":
                b.set_mlp(params)
            b.sync()
            t0 = time.perf_counter()
            b.step_async(steps, use_graph=True)
            b.sync()
            el = time.perf_counter() - t0
            s = b.stats()
            b.close()
            row = {"iterations": it, "warm_start": mode,
                   "body_steps_per_s": round(sc.n_dynamic * steps / el, 1),
                   "max_penetration": s["run_max_depth"],
                   "normal_residual": s["run_normal_residual"],
                   "friction_residual": s["run_friction_residual"],
                   "mean_step_max_penetration": s["mean_max_depth"],
                   "mean_step_normal_residual": s["mean_normal_residual"],
                   "mean_step_friction_residual": s["mean_friction_residual"]}
            log(json.dumps(row))
            out.append(row)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--samples", type=int, default=5_000_000)
    c.add_argument("--count", type=int, default=8192)
    c.add_argument("--steps", type=int, default=600)
    c.add_argument("--scene", default="polygon_pile", choices=["polygon_pile", "pile_batch"])
    c.add_argument("--worlds", type=int, default=64)
    c.add_argument("--skip", type=int, default=0)
    c.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--data", required=True)
    t.add_argument("--out", default=DEFAULT_MODEL)
    t.add_argument("--epochs", type=int, default=8)
    t.add_argument("--batch", type=int, default=4096)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--huber", type=float, default=0.0)
    t.add_argument("--clip", type=float, default=20.0)
    e = sub.add_parser("eval")
    e.add_argument("--model", default=DEFAULT_MODEL)
    e.add_argument("--steps", type=int, default=600)
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    if a.cmd == "collect":
        f, tg = collect(a.samples, count=a.count, steps=a.steps, scene=a.scene, worlds=a.worlds,
                        skip=a.skip, log=log)
        np.savez(a.out, feats=f, targets=tg)
        log(f"collect: wrote {len(f)} samples to {a.out}")
    elif a.cmd == "train":
        d = np.load(a.data)
        p, hist = train(d["feats"], d["targets"], epochs=a.epochs, batch=a.batch, lr=a.lr,
                        huber=a.huber, clip=a.clip, log=log)
        np.save(a.out, p)
        log(f"train: wrote {a.out}")
    else:
        p = np.load(a.model)
        evaluate(p, steps=a.steps, log=log)


if __name__ == "__main__":
    sys.exit(main())
