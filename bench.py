"""Benchmark of the contact-physics step on one MI355X.

Headline (no flags): body-steps/s at fixed penetration/residual, batched +
learned warm start against the plain GPU solver, on the combined-at-scale
workload: 1024 independent worlds of 2k-32k random 3-8-gons stacked in
piles up to 60 layers (11.3 M bodies), coloured FP32 solver, dt 1/120, 4
sweeps (scenes.pile_batch).  A "step" is one updateWorld of every world of
the batch; body-steps/s counts dynamic bodies.

Every run creates a batch from the initial state, steps it through an
untimed pre-roll (`--warmup` plus the workload's settling steps, half of its
configured run) so that the timed window sits in the settled regime
whatever `--steps` is, then times exactly `--steps` graph-replayed steps
with inputs resident in HBM.  Quality is compared over the whole run
(pre-roll and window): the run means of the per-step maximum penetration
and normal residual.

1. autotune: broadphase x solver lanes per world, timed in the settled
   regime (plain solver);
2. plain GPU: zero start at the workload's sweeps;
3. learned warm start at the same and then fewer sweeps while its quality
   stays within 2% of the plain run's: `value` is the body-steps/s of the
   fewest sweeps that match, or the plain run's when none does
   (`headline.value_is` says which);
4. the FP64 drop-in solver over the same steps: quality and the divergence
   of each run's final positions, beside the FP64 run's own chaos floor;
5. the learned warm start on the 8192-polygon world at 2 and 4 sweeps
   against plain at 8 (`c4_learned_warm_start`);
6. an instrumented replay (eager launches, HIP events on the stream each
   kernel runs on) of the plain configuration after its window: per-phase
   times and the roofline of the dominant kernel;
7. the CPU oracle on this host over the same step window on a sample of
   the worlds (`cpu_baseline`).

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from shapes_amd import engine, scenes  # noqa: E402

METRIC = "body-steps/s at fixed penetration/residual: batched+learned vs plain GPU"
HBM_PEAK_GBS = 8000.0
QUALITY_TOL = 0.02

SOLVERS = {"ordered": engine.SOLVER_ORDERED_F64, "colored": engine.SOLVER_COLORED_F32}
DTYPES = {"ordered": "f64", "colored": "f32"}
BROADPHASES = {"auto": engine.BROADPHASE_AUTO, "scan": engine.BROADPHASE_SCAN,
               "grid": engine.BROADPHASE_GRID, "grid1": engine.BROADPHASE_GRID_1,
               "grid16": engine.BROADPHASE_GRID_16}

# workload -> (scene factory, settling pre-roll, pair capacity per body,
# autotune candidates, description).  The pre-rolls end where the contact
# count has settled: the 1024 piles collapse and rebound until about step
# 700 (contacts 15-28 M, 22-32 ms per step) and then hold 22.6-23.6 M
# contacts; the 8192-polygon pile lands around step 330 and settles by 500;
# the 4096-world batch settles by 170.
WORKLOADS = {
    "pile_batch": (lambda w: scenes.pile_batch(n_worlds=w or 1024), 800, 4,
                   [("grid1", 512), ("grid1", 1024)],
                   "worlds of 2k-32k random 3-8-gons stacked in piles up to 60 layers"),
    "world_batch": (lambda w: scenes.world_batch_fast(n_worlds=w or 4096), 150, 4,
                    [(b, t) for b in ("grid1", "grid16", "scan") for t in (64, 256)] +
                    [("grid1", 512), ("grid1", 1024)],
                    "worlds of 8-1024 bodies (squares and 3-8-gons), each on its own floor"),
    "polygon_pile": (lambda w: scenes.polygon_pile(), 500, 0,
                     [(b, t) for b in ("grid", "scan") for t in (256, 512, 1024)],
                     "one world of 8192 random 3-8-gons dropped onto a 200-wide floor"),
    "pyramid": (lambda w: scenes.pyramid(), 150, 0, [("scan", 256)],
                "pyramid of 55 unit boxes on a static floor"),
}


def note(msg: str) -> None:
    """Progress of a long run on stderr (stdout carries the JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def quality(st):
    return {"max_penetration": st["run_max_depth"],
            "mean_step_max_penetration": st["mean_max_depth"],
            "mean_step_normal_residual": st["mean_normal_residual"],
            "mean_step_friction_residual": st["mean_friction_residual"],
            "normal_residual": st["run_normal_residual"],
            "friction_residual": st["run_friction_residual"]}


def matches(q, ref, tol=QUALITY_TOL):
    """Run means of the per-step maximum penetration and normal residual
    within `tol` of the reference run's."""
    return (q["mean_step_max_penetration"] <= ref["mean_step_max_penetration"] * (1 + tol) and
            q["mean_step_normal_residual"] <= ref["mean_step_normal_residual"] * (1 + tol))


GRAPH = True  # steps replayed from a captured HIP graph (--no-graph: eager launches)


def run(scene, solver, warm, iterations, preroll, steps, sched, ppb, model=None,
        positions=False, instrument=0):
    """One batch from the initial state: `preroll` untimed graph-replayed
    steps, then exactly `steps` timed ones; optionally `instrument` eager
    steps with HIP-event timing after them.  Returns a dict with the timed
    seconds, the run's stats and quality, final positions (at the end of
    the timed window) and the instrumented phase times."""
    cfg = scenes.SceneConfig(**{**scene.config.__dict__, "iterations": iterations})
    b = engine.make_batch(scene.worlds, cfg, solver=solver, warm_start=warm,
                          broadphase=BROADPHASES[sched[0]], max_pairs_per_body=ppb,
                          solver_threads=sched[1])
    try:
        if warm == engine.WARM_LEARNED:
            b.set_mlp(model)
        if preroll:
            b.step_async(preroll, use_graph=GRAPH)
        b.sync()
        t0 = time.perf_counter()
        b.step_async(steps, use_graph=GRAPH)
        b.sync()
        elapsed = time.perf_counter() - t0
        st = b.stats()
        out = {"seconds": elapsed, "stats": st, "quality": quality(st)}
        if positions:
            out["positions"] = [b.bodies(k)[:, :2].copy() for k in range(len(scene.worlds))]
        if instrument:
            b.set_timing(True)
            tot = {"pairs": 0, "contacts": 0, "manifolds": 0}
            for _ in range(instrument):
                b.step(1)
                s = b.stats()
                tot["pairs"] += s["pairs"]
                tot["contacts"] += s["contacts"]
                tot["manifolds"] += s["manifolds"]
            t = b.timing()
            out["phases_ms"] = {p: t[p] for p in engine.PHASES}
            out["solve_profile_us"] = b.solve_profile()
            out["instrumented"] = {k: v / instrument for k, v in tot.items()}
            out["instrumented"]["steps"] = instrument
    finally:
        b.close()
    return out


def autotune(scene, ppb, candidates, preroll, steps=30, solver=engine.SOLVER_COLORED_F32):
    """Pick broadphase x solver lanes per world by timing `steps` graph
    replays of the plain solver after `preroll` untimed steps (the ordered
    solver sizes its lanes by world size: only the broadphase is tuned)."""
    if solver == engine.SOLVER_ORDERED_F64:
        candidates = sorted({(b, 0) for b, _ in candidates})
    times = {}
    for sched in candidates:
        r = run(scene, solver, engine.WARM_ZERO, scene.config.iterations,
                preroll, steps, sched, ppb)
        times[sched] = r["seconds"] / steps * 1e3
        note(f"autotune {sched[0]}/{sched[1]}: {times[sched]:.3f} ms/step")
    best = min(times, key=times.get)
    return best, {f"{k[0]}/{k[1]}": round(v, 4) for k, v in times.items()}


def divergence(pa, pb, worlds):
    """Distance between two runs' final positions over the dynamic bodies."""
    d = np.concatenate([np.hypot(a[:, 0] - b[:, 0], a[:, 1] - b[:, 1])[(w.mass != 0) | (w.inertia != 0)]
                        for a, b, w in zip(pa, pb, worlds)])
    return {"max": round(float(d.max()), 6), "rms": round(float(np.sqrt((d ** 2).mean())), 6)}


# --------------------------------------------------------------- roofline

def kernel_bytes(phase, nb, nv, st, word):
    """Algorithmic (compulsory) HBM bytes of one launch of a phase's kernel,
    from the step's counts: nb bodies, nv hull vertices, st['pairs'],
    st['contacts'] and st['manifolds'] (pairs with stored contacts)."""
    P, C, M = st["pairs"], st["contacts"], st["manifolds"]
    if phase == "sweeps":
        # per manifold its record (normal, materials, normal x inverse
        # masses, bodies, flags, first contact's 11 FP32 fields and two FP64
        # lagrangians: 112 B) read and the lagrangians written (16 B); per
        # second contact its record (64 B) read and lagrangians written
        # (16 B); per body the FP64 velocity read (24 B) and the FP32
        # velocity written (12 B)
        return M * (112 + 16) + (C - M) * (64 + 16) + nb * (24 + 12)
    if phase == "colouring":
        # per pair its body word and contact count read, colour and slot
        # written (16 B); per body its colour mask and priority (12 B)
        return P * 16 + nb * 12
    if phase == "finish":
        # advance + transform (FP64 in both modes): state read (6 doubles)
        # and pose written (3), AABB written (4); per vertex the local vertex
        # and extent indices read (18 B), world vertex, normal and own
        # extent written (48 B); quality: per contact bodies, velocities,
        # cross terms, normal, flip, radii, bias, lagrangians, mass,
        # depth, materials (209 B)
        return nb * (13 * 8 + 4) + nv * (18 + 48) + C * 209
    if phase == "sat":
        # per pair its two body indices read, its contact count and (coloured)
        # its colouring word written; every hull's world vertices, normals
        # and own extents read once (FP64 in both modes); per contact slot
        # its normal, point, depth, feature and flip written
        return P * (8 + 4 + (4 if word == 4 else 0)) + nv * 48 + C * 43
    if phase == "contacts":
        # per contact its slot (43 B) and the cache entry it joins (24 B)
        # read, and key, bodies, flip, normal, point, depth, the rows' cross
        # terms, bias, radii, effective masses, lagrangians and warm flag
        # written (162 B); per pair with contacts both bodies' positions and
        # inverse masses read (64 B); for the coloured solver each manifold's
        # record written (112 B, 64 B per second contact) with both bodies'
        # materials read (34 B)
        rec = M * (112 + 34) + (C - M) * 64 if word == 4 else 0
        return C * (43 + 24 + 162) + M * 64 + rec
    if phase == "solve_writeback":  # lagrangians back, FP32 -> FP64 velocities
        return C * (16 + 16) + nb * (12 + 48)
    if phase == "broadphase_fill":
        return nb * (48 + 32 + 8) + P * 8
    if phase == "infer":
        return C * 60
    if phase == "cache_store":
        return C * 48
    return 0.0


def roofline(instr, scene, word, broadphase="grid", threads=0):
    """The dominant kernel of the instrumented replay (the phase with the
    largest average HIP-event time among those holding a single kernel):
    algorithmic bytes per launch / its average launch duration."""
    ph = instr["phases_ms"]
    single = ["sat", "contacts", "infer", "sweeps", "solve_writeback", "finish",
              "cache_store", "colouring", "broadphase_fill"]
    dom = max(single, key=lambda p: ph[p])
    ms = ph[dom]
    nb = scene.n_bodies
    nv = sum(int(w.hull_off[-1]) for w in scene.worlds)
    by = kernel_bytes(dom, nb, nv, instr["instrumented"], word)
    achieved = by / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {
        "bound": "hbm",
        "kernel": phase_kernel(dom, word, broadphase, threads,
                               max(w.n for w in scene.worlds)),
        "phase": dom,
        "achieved": round(achieved, 3),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 6),
        "traffic": None,
        "algorithmic_bytes_per_launch": round(by),
        "avg_launch_us": round(ms * 1e3, 2),
        "launches_timed": instr["instrumented"]["steps"],
    }


# worlds of at most this many bodies sweep one wavefront each when the
# coloured solver runs with 64 lanes
SWEEP_WAVE_MAX_BODIES = 1024


def phase_kernel(phase, word, broadphase, threads=0, largest=0):
    """The kernel a single-kernel phase launches in this configuration (the
    coloured sweeps with 64 lanes: k_sweep_waves alone when every world
    fits a wavefront's slice, else it and k_sweep_colours)."""
    if phase == "sweeps":
        if word == 8:
            return "k_solve_flow"
        if threads != 64:
            return "k_sweep_colours"
        return ("k_sweep_waves" if largest <= SWEEP_WAVE_MAX_BODIES
                else "k_sweep_colours+k_sweep_waves")
    if phase == "broadphase_fill":
        return "k_broadphase" if broadphase == "scan" else "k_grid_pairs"
    return engine.PHASE_KERNELS[phase][0]


def step_bytes(nb, nv, contacts, word):
    """Compulsory HBM bytes of one whole step: every body's state and
    materials read, pos/theta/vel/omega written ((10 + 6) words), its local
    hull read (2 words per vertex), and the warm-start cache read and
    written (packed key + two FP64 lambdas = 24 B each way per contact)."""
    return nb * 16 * word + nv * 2 * word + contacts * 48


# ------------------------------------------------------------ CPU baseline

def cpu_sample(scene, max_worlds=16, max_bodies=8000):
    """World indices of the CPU sample: one world, or up to `max_worlds`
    worlds spread over the batch's size ranks among those of at most
    `max_bodies` bodies (the oracle steps each world on one thread)."""
    if len(scene.worlds) == 1:
        return [0]
    sizes = np.array([w.n for w in scene.worlds])
    cand = np.nonzero(sizes <= max_bodies)[0]
    cand = cand[np.argsort(sizes[cand], kind="stable")]
    pick = np.unique(np.linspace(0, len(cand) - 1, min(max_worlds, len(cand))).astype(int))
    return sorted(cand[pick].tolist())


def cpu_baseline(scene, preroll, steps):
    """The C++ FP64 oracle on this host over the same step window as the GPU
    run: the sampled worlds step `preroll` untimed steps (all threads), then
    `steps` timed steps on one thread and on all threads (up to 16)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    oracle.build()
    idx = cpu_sample(scene)
    worlds = [scene.worlds[k] for k in idx]
    threads = max(1, min(16, os.cpu_count() or 1, len(worlds)))
    engines = [oracle.OracleEngine(w, scene.config, world_index=k) for w, k in zip(worlds, idx)]
    try:
        pre = oracle.run_engines(engines, preroll, threads=threads) if preroll else 0.0
        one = oracle.run_engines(engines, steps, threads=1)
        many = oracle.run_engines(engines, steps, threads=threads) if threads > 1 else one
    finally:
        for e in engines:
            e.close()
    dyn = sum(w.n_dynamic for w in worlds)
    what = (f"{scene.name} ({dyn} dynamic bodies)" if len(worlds) == 1 else
            f"{len(worlds)} of the {len(scene.worlds)} worlds of {scene.name} ({dyn} dynamic "
            f"bodies; indices {idx})")
    return {
        "value": round(dyn * steps / many, 1),
        "unit": "body-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"steps {preroll}-{preroll + steps} (after an untimed pre-roll of {preroll} steps, "
                  f"{pre:.1f} s) of {what}, C++ FP64 oracle with the reference's grid "
                  f"broadphase, {threads} thread(s): {many:.2f} s; the same window on 1 thread: "
                  f"{one:.2f} s",
        "single_thread_value": round(dyn * steps / one, 1),
    }


# ------------------------------------------------------------------ parts

def headline(scene, args, model, sched, ppb, dyn, preroll):
    """Plain coloured solver at the workload's sweeps, then the learned
    warm start at the same and fewer sweeps while its quality matches."""
    base = scene.config.iterations
    col = engine.SOLVER_COLORED_F32
    r = run(scene, col, engine.WARM_ZERO, base, preroll, args.steps, sched, ppb,
            positions=True, instrument=args.profile_steps)
    plain = {"warm_start": "zero", "iterations": base,
             "body_steps_per_s": round(dyn * args.steps / r["seconds"], 1),
             "ms_per_step": round(1e3 * r["seconds"] / args.steps, 4), **r["quality"]}
    note(f"plain coloured at {base} sweeps: {plain['ms_per_step']} ms/step")
    runs = {("zero", base): r}
    tried, chosen = [], None
    if model is not None:
        for it in range(base, 0, -1):
            try:
                rl = run(scene, col, engine.WARM_LEARNED, it, preroll, args.steps, sched, ppb,
                         model=model, positions=True, instrument=args.profile_steps)
            except engine.ShapesError as e:
                tried.append({"warm_start": "learned", "iterations": it, "error": str(e)})
                note(f"learned at {it} sweeps: {e}")
                break
            row = {"warm_start": "learned", "iterations": it,
                   "body_steps_per_s": round(dyn * args.steps / rl["seconds"], 1),
                   "ms_per_step": round(1e3 * rl["seconds"] / args.steps, 4), **rl["quality"],
                   "learned_starts_last_step": rl["stats"]["learned_starts"],
                   "misses_last_step": rl["stats"]["misses"]}
            row["matches_plain"] = matches(row, plain)
            tried.append(row)
            runs[("learned", it)] = rl
            note(f"learned at {it} sweeps: {row['ms_per_step']} ms/step, matches plain: "
                 f"{row['matches_plain']}")
            if row["matches_plain"]:
                chosen = row
            elif it < base:
                break
    return plain, tried, chosen, runs


def fp64_reference(scene, args, sched, ppb, preroll, runs):
    """The FP64 drop-in solver (the reference's order, bit-exact to the
    oracle) over the same steps: its quality, and the divergence of each
    run's final positions from it; `chaos_floor` is the FP64 run's own
    divergence when every dynamic body's initial x-velocity moves by 1e-12."""
    base = scene.config.iterations
    total = preroll + args.steps
    ordered = engine.SOLVER_ORDERED_F64
    sched = (sched[0], 0)  # the ordered solver's lanes by world size
    ref = run(scene, ordered, engine.WARM_ZERO, base, 0, total, sched, ppb, positions=True)
    moved = []
    for w in scene.worlds:
        m = w.copy()
        m.vx = m.vx + np.where((m.mass != 0) | (m.inertia != 0), 1e-12, 0.0)
        moved.append(m)
    msc = scenes.Scene(moved, scene.config, name=scene.name)
    pert = run(msc, ordered, engine.WARM_ZERO, base, 0, total, sched, ppb, positions=True)
    out = {"solver": "ordered", "iterations": base, "steps": total,
           "ms_per_step": round(1e3 * ref["seconds"] / total, 4),
           "quality": ref["quality"],
           "chaos_floor": divergence(pert["positions"], ref["positions"], scene.worlds),
           "runs": []}
    for (warm, it), r in runs.items():
        out["runs"].append({"warm_start": warm, "iterations": it,
                            "divergence": divergence(r["positions"], ref["positions"],
                                                     scene.worlds),
                            "mean_step_max_penetration": r["quality"]["mean_step_max_penetration"],
                            "mean_step_normal_residual": r["quality"]["mean_step_normal_residual"]})
    return out


def c4_evaluation(model, seeds=(1002, 1003, 1004)):
    """The learned warm start on the 8192-polygon world (dt 1/60, 600
    steps) at 2 and 4 sweeps against the plain solver at 8, quality averaged
    over three seeds of the generator (not the model's training seeds),
    throughput of the first."""
    rows = []
    for warm, it in ((engine.WARM_ZERO, 8), (engine.WARM_LEARNED, 4), (engine.WARM_LEARNED, 2),
                     (engine.WARM_ZERO, 4), (engine.WARM_ZERO, 2)):
        qs, secs = [], None
        for sd in seeds:
            sc = scenes.polygon_pile(seed=sd)
            r = run(sc, engine.SOLVER_COLORED_F32, warm, it, 0, sc.config.steps, ("grid", 1024),
                    0, model=model)
            qs.append(r["quality"])
            if secs is None:
                secs = r["seconds"]
        keys = ("mean_step_max_penetration", "mean_step_normal_residual",
                "mean_step_friction_residual")
        rows.append({"warm_start": "learned" if warm == engine.WARM_LEARNED else "zero",
                     "iterations": it, **{k: float(np.mean([q[k] for q in qs])) for k in keys},
                     "ms_per_step": round(1e3 * secs / 600, 4)})
    ref = rows[0]
    for r in rows[1:]:
        r["matches_plain_8"] = matches(r, ref)
    return {"workload": "polygon_pile8192, 600 steps, seeds " + ", ".join(map(str, seeds)),
            "rows": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed steps (K)")
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed steps before the settling pre-roll (W)")
    ap.add_argument("--workload", default="pile_batch", choices=list(WORKLOADS))
    ap.add_argument("--worlds", type=int, default=0, help="batch workloads: number of worlds")
    ap.add_argument("--mode", default="headline", choices=["headline", "single"],
                    help="headline: learned vs plain at fixed quality; single: one configuration")
    ap.add_argument("--solver", default="colored", choices=list(SOLVERS))
    ap.add_argument("--warm", default="zero", choices=["zero", "learned"])
    ap.add_argument("--iterations", type=int, default=0, help="sweeps (default: the workload's)")
    ap.add_argument("--preroll", type=int, default=-1,
                    help="settling steps before the window (default: half the workload's run)")
    ap.add_argument("--model", default=os.path.join(ROOT, "shapes_amd", "warmstart_mlp.npy"))
    ap.add_argument("--profile-steps", type=int, default=20,
                    help="instrumented eager steps after the plain run's window")
    ap.add_argument("--broadphase", default="tune", choices=["tune"] + list(BROADPHASES))
    ap.add_argument("--solver-threads", type=int, default=0, choices=[0, 64, 256, 512, 1024])
    ap.add_argument("--pairs-per-body", type=int, default=-1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="eager launches instead of graph replays (the same kernels; for "
                         "profilers)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device three times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        x = torch.empty(1 << 30, dtype=torch.uint8, device="cuda")
        x.fill_(1)
        for _ in range(3):
            y = x.clone()
            del y
        torch.cuda.synchronize()
        print("copied 1073741824 bytes x3", flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build steps worlds on one GPU only; sharding worlds "
              "or slabs over several GPUs is not implemented (run with --gpus 1).",
              file=sys.stderr)
        sys.exit(2)

    global GRAPH
    GRAPH = not args.no_graph
    t_start = time.perf_counter()
    make, settle, ppb_default, candidates, desc = WORKLOADS[args.workload]
    scene = make(args.worlds)
    note(f"scene {scene.name}: {scene.n_bodies} bodies in {len(scene.worlds)} world(s), "
         f"built in {time.perf_counter() - t_start:.1f} s")
    if args.iterations:
        scene.config.iterations = args.iterations
    preroll = args.warmup + (settle if args.preroll < 0 else args.preroll)
    ppb = ppb_default if args.pairs_per_body < 0 else args.pairs_per_body
    dyn = scene.n_dynamic
    headline_mode = args.mode == "headline"
    solver_name = "colored" if headline_mode else args.solver
    model = None
    if headline_mode or args.warm == "learned":
        if not os.path.exists(args.model):
            raise SystemExit(f"bench.py: no model at {args.model}")
        model = np.load(args.model)

    tuning = None
    if args.broadphase == "tune":
        sched, tuning = autotune(scene, ppb, candidates, preroll, solver=SOLVERS[solver_name])
    else:
        sched = (args.broadphase, args.solver_threads)
    note(f"schedule: {sched[0]}, {sched[1]} lanes")

    extra = {}
    if headline_mode:
        plain, tried, chosen, runs = headline(scene, args, model, sched, ppb, dyn, preroll)
        r_plain = runs[("zero", scene.config.iterations)]
        if chosen is not None:
            value = chosen["body_steps_per_s"]
            ms = chosen["ms_per_step"]
            it = chosen["iterations"]
            warm_name = "learned"
            value_is = f"learned warm start at {it} sweeps (quality within 2% of plain)"
            q = {k: chosen[k] for k in plain if k.startswith(("max_", "mean_", "normal_",
                                                              "friction_"))}
        else:
            value = plain["body_steps_per_s"]
            ms = plain["ms_per_step"]
            it = plain["iterations"]
            warm_name = "zero"
            value_is = ("plain GPU solver: no learned configuration matched its quality"
                        if model is not None else "plain GPU solver")
            q = r_plain["quality"]
        extra["headline"] = {
            "value_is": value_is,
            "model": os.path.relpath(args.model, ROOT),
            "quality_matched": chosen is not None,
            "quality_tolerance": QUALITY_TOL,
            "plain_gpu": plain,
            "learned_runs": tried,
            "vs_plain_gpu": round(value / plain["body_steps_per_s"], 4),
            "sweeps_plain_over_learned": (round(plain["iterations"] / chosen["iterations"], 3)
                                          if chosen else None),
        }
        # phases and roofline of the configuration `value` is quoted on
        instr = runs[("learned", chosen["iterations"])] if chosen is not None else r_plain
        if not args.no_fp64:
            extra["fp64_reference"] = fp64_reference(scene, args, sched, ppb, preroll, runs)
            note("FP64 reference runs done")
        if not args.no_c4 and model is not None:
            extra["c4_learned_warm_start"] = c4_evaluation(model)
            note("learned warm start on the 8192-polygon world done")
    else:
        warm = engine.WARM_LEARNED if args.warm == "learned" else engine.WARM_ZERO
        r = run(scene, SOLVERS[solver_name], warm, scene.config.iterations, preroll, args.steps,
                sched, ppb, model=model, instrument=args.profile_steps)
        value = dyn * args.steps / r["seconds"]
        ms = 1e3 * r["seconds"] / args.steps
        it = scene.config.iterations
        warm_name = args.warm
        q = r["quality"]
        instr = r
        st = r["stats"]
        extra["last_step"] = {k: st[k] for k in ("pairs", "contacts", "warm_hits", "misses",
                                                 "learned_starts", "max_levels")}

    word = 4 if solver_name == "colored" else 8
    nv = sum(int(w.hull_off[-1]) for w in scene.worlds)
    roof = roofline(instr, scene, word, sched[0], sched[1])
    sb = step_bytes(scene.n_bodies, nv, instr["instrumented"]["contacts"], word)
    step_gbs = sb / (ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[solver_name],
        "data": "synthetic",
        "config": {
            "workload": f"{scene.name}: {len(scene.worlds)} {desc}" if len(scene.worlds) > 1
            else f"{scene.name}: {desc}",
            "worlds": len(scene.worlds),
            "bodies": scene.n_bodies,
            "dynamic_bodies": dyn,
            "iterations": it,
            "dt": scene.config.dt,
            "solver": solver_name,
            "warm_start": warm_name,
            "mode": args.mode,
            "graph": GRAPH,
            "timed_window": [preroll, preroll + args.steps],
            "broadphase": sched[0],
            "solver_threads": sched[1],
            "pairs_per_body": ppb,
            "autotune_ms_per_step": tuning,
        },
        "quality": q,
        **extra,
        "roofline": roof,
        "step_roofline": {"bound": "hbm", "algorithmic_bytes_per_step": round(sb),
                          "achieved": round(step_gbs, 3), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 6)},
        "phase_us": {p: round(v * 1e3, 2) for p, v in instr["phases_ms"].items()},
        "largest_world_solve_us": instr["solve_profile_us"],
        "avg_pairs": round(instr["instrumented"]["pairs"], 1),
        "avg_contacts": round(instr["instrumented"]["contacts"], 1),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, preroll, min(args.steps, 20))
        note("cpu baseline done")
    note(f"total {time.perf_counter() - t_start:.0f} s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
