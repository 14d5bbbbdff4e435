"""Upper bound of any warm start for new contacts, on the CPU oracle.

A learned warm start can only change the starting impulses of contacts
that missed the cache (Solvers/Contact.hs:91-99, the zero start of
newCache).  The best any predictor can do is to start them at the
impulses the solve itself ends with.  This experiment runs, every step,
an 8-sweep solve from the current state to obtain those impulses, then
the k-sweep step whose new contacts start from them ("ideal@k"), and
compares the solve quality with the zero start at k and 8 sweeps.

Quality per step: maxima over contacts of penetration depth and the
normal / friction residuals; summarised over the run by their means.

    python tests/test_warmstart_bound.py --steps 600    # the full pile

The CPU test runs a small pile for a few dozen steps and checks the
machinery: restarting from a copied state reproduces the plain run bit for
bit, and the ideal start at 2 sweeps lowers the mean normal residual.
"""
import argparse
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from shapes_amd import scenes  # noqa: E402


def _orc():
    import oracle
    oracle.build()
    return oracle


def world_from(w0, S):
    """w0's bodies at state S (n x 6: px, py, theta, vx, vy, omega)."""
    return scenes.WorldDesc(hull_off=w0.hull_off, hull_x=w0.hull_x, hull_y=w0.hull_y,
                            px=S[:, 0].copy(), py=S[:, 1].copy(), theta=S[:, 2].copy(),
                            vx=S[:, 3].copy(), vy=S[:, 4].copy(), omega=S[:, 5].copy(),
                            mass=w0.mass, inertia=w0.inertia, mu=w0.mu, bounce=w0.bounce)


def _q(e):
    s = e.stats()
    return np.array([s["max_depth"], s["normal_residual"], s["friction_residual"]])


def run(orc, scene, mode, iters, steps, restart=False):
    """Per-step quality rows and the final state of one run.  mode 'plain':
    zero start for new contacts; 'ideal': new contacts start from the
    impulses an 8-sweep solve of the same step ends with."""
    w0 = scene.worlds[0]
    cfg = scenes.SceneConfig(dt=scene.config.dt, gy=scene.config.gy, iterations=iters)
    cfg8 = scenes.SceneConfig(dt=scene.config.dt, gy=scene.config.gy, iterations=8)
    rows = []
    if mode == "plain" and not restart:
        e = orc.OracleEngine(w0, cfg)
        for _ in range(steps):
            e.step(1)
            rows.append(_q(e))
        return np.array(rows), e.bodies()
    S = orc.OracleEngine(w0, cfg).bodies()
    keys, lam = np.zeros((0, 4), np.int32), np.zeros((0, 2))
    for _ in range(steps):
        w = world_from(w0, S)
        if mode == "ideal":
            e8 = orc.OracleEngine(w, cfg8)
            e8.set_cache(keys, lam)
            e8.step(1)
            c8 = e8.contacts()
            e8.close()
            new = c8["warm"] == 0
            k2 = np.concatenate([keys, c8["keys"][new]])
            l2 = np.concatenate([lam, c8["lam"][new]])
            order = np.lexsort((k2[:, 3], k2[:, 2], k2[:, 1], k2[:, 0]))[::-1]
            keys, lam = k2[order], l2[order]   # the cache is kept descending
        e = orc.OracleEngine(w, cfg)
        e.set_cache(keys, lam)
        e.step(1)
        rows.append(_q(e))
        c = e.contacts()
        S = e.bodies()
        keys, lam = c["keys"], c["lam"]
        e.close()
    return np.array(rows), S


def summary(rows):
    return {"max": rows.max(axis=0).tolist(), "mean": rows.mean(axis=0).tolist()}


def test_restart_reproduces_plain_and_ideal_helps():
    orc = _orc()
    sc = scenes.polygon_pile(count=1024, steps=40)
    plain, s1 = run(orc, sc, "plain", 2, 40)
    again, s2 = run(orc, sc, "plain", 2, 40, restart=True)
    assert np.array_equal(s1.view(np.uint64), s2.view(np.uint64))
    assert np.array_equal(plain, again)
    ideal, _ = run(orc, sc, "ideal", 2, 40)
    assert ideal[:, 1].mean() < plain[:, 1].mean()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--count", type=int, default=8192)
    args = ap.parse_args()
    orc = _orc()
    sc = scenes.polygon_pile(count=args.count, steps=args.steps)
    out = {}
    for mode, it in [("plain", 8), ("plain", 6), ("plain", 4), ("ideal", 4), ("plain", 2),
                     ("ideal", 2)]:
        rows, _ = run(orc, sc, mode, it, args.steps)
        out[f"{mode}@{it}"] = summary(rows)
        print(f"{mode}@{it}", json.dumps(out[f"{mode}@{it}"]), flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
