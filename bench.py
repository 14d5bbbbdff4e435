"""Benchmark of the contact-physics step on one MI355X.

Default workload: one large world of 8192 random convex 3-8-gons dropped
onto a 200-wide static floor, dt = 1/60, 600 steps (scenes.polygon_pile).
A "step" is one updateWorld of that world; body-steps/s counts dynamic
bodies.  Each timed region replays K steps from the initial state (as HIP
graphs) after W warm-up steps on a separate instance of the same scene.

Default mode ("headline"): body-steps/s at fixed penetration/residual.  The
plain GPU solver (coloured, zero start for new contacts, the workload's 8
sweeps) sets the quality: the run means of the per-step maximum
penetration and normal residual.  The learned warm start then runs at 8,
7, 6, ... sweeps while both means stay within 2% of the plain run's; the
value is the body-steps/s of the fewest sweeps that pass (at least the
workload's sweep count when none below it does), next to the plain run's.

"--mode single" times one configuration (--solver, --warm, --iterations).

Prints ONE JSON line: the value, the roofline of the dominant kernel
(algorithmic bytes per launch / its average duration, measured with HIP
events on the batch's stream in an instrumented replay of the same steps),
the solve quality, and the CPU oracle timed on this host on a bounded
sample of the same workload.
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

SOLVERS = {"ordered": engine.SOLVER_ORDERED_F64, "colored": engine.SOLVER_COLORED_F32}
DTYPES = {"ordered": "f64", "colored": "f32"}
DTYPES_BYTES = {"ordered": 8, "colored": 4}


def make_scene(name: str, worlds: int = 0):
    if name == "polygon_pile":
        return scenes.polygon_pile()
    if name == "world_batch":
        return scenes.world_batch_fast(n_worlds=worlds or 4096)
    if name == "pile_batch":
        return scenes.pile_batch(n_worlds=worlds or 1024)
    if name == "pyramid":
        return scenes.pyramid()
    if name == "reference_stacks":
        return scenes.reference_stacks()
    raise SystemExit(f"unknown workload {name}")


# --------------------------------------------------------------- roofline

# the kernels of a phase (its HIP-event time covers them; the coloured
# solve's k_colour runs on a second stream from the end of k_sat, and the
# phase waits for it)
PHASE_KERNELS = {
    ("solve", "colored"): ["k_colour", "k_pack_manifolds", "k_sweep_colours", "k_colour_unpack"],
    ("solve", "ordered"): ["k_solve_flow"],
    ("sat", None): ["k_sat"],
    ("contacts", None): ["k_contacts"],
    # advance and the solve quality share one launch (in the quality phase)
    ("quality", None): ["k_finish"],
    ("integrate", None): ["k_finish"],
    ("broadphase_count", None): ["k_grid_assign", "k_grid_fill", "k_grid_pairs", "k_broadphase"],
    ("broadphase_fill", None): ["k_grid_pairs", "k_broadphase"],
}


def kernel_bytes(phase: str, st: dict, nb: int, nv: int, word: int) -> float:
    """Algorithmic (compulsory) HBM bytes of one launch of a phase, from
    the step's counts: nb bodies, nv hull vertices, st['pairs'],
    st['contacts'] and the solver's passes.  FP64 words are 8 bytes."""
    P, C = st["pairs"], st["contacts"]
    passes = st["passes"]
    if phase == "integrate":
        # read pos, theta, vel, omega (6W) + write pos, theta (3W) + vert
        # offsets; read local verts (2W/vertex), write world verts and
        # normals (4W/vertex) and the AABB (4W)
        return nb * (9 * word + 4 + 4 * word) + nv * 6 * word
    if phase in ("broadphase_count", "broadphase_fill"):
        out = 4 * nb if phase == "broadphase_count" else 8 * P
        return nb * (4 * word + 1 + 8) + out
    if phase == "sat":
        # pair ids, both hulls' verts + normals + extents, <= 2 slots out
        return P * (8 + 2 * 8) + 2 * (nv / max(nb, 1)) * P * (4 * word + 2) + C * 50 + 4 * P
    if phase == "contacts":
        # slot in, body positions and inverse masses, cache lookup; out:
        # key, bodies, flip, geometry, two 6-rows, bias, radii, masses, lambdas
        return C * (50 + 2 * 4 * word + 24) + C * (8 + 8 + 1 + 5 * word + 12 * word +
                                                   word + 4 * word + 2 * word + 2 * 8 + 1)
    if phase == "solve":
        # compulsory traffic of one launch: per contact its bodies and warm
        # flag, the normal, the four cross terms, bias, lever arms, two
        # effective masses and two lambdas read (15 FP64 words) and two
        # lambdas written; per body the velocity read and written and the
        # inverse masses and materials read.  Re-reads of the same records
        # and velocities across sweeps are served by LDS and L2.
        return C * (9 + 15 * 8 + 2 * 8) + nb * (6 * 8 + 4 * 8 + 1)
    return 0.0


def instrumented_run(scene, solver, steps, broadphase="auto", ppb=0, warm=engine.WARM_ZERO,
                     model=None, threads=0):
    w = scene.worlds
    b = engine.make_batch(w, scene.config, solver=solver, broadphase=BROADPHASES[broadphase],
                          max_pairs_per_body=ppb, warm_start=warm, solver_threads=threads)
    if warm == engine.WARM_LEARNED:
        b.set_mlp(model)
    b.set_timing(True)
    tot = {"pairs": 0, "contacts": 0}
    for _ in range(steps):
        b.step(1)
        s = b.stats()
        tot["pairs"] += s["pairs"]
        tot["contacts"] += s["contacts"]
    t = b.timing()
    t["solve_profile_us"] = b.solve_profile()
    b.close()
    n = max(t["steps"], 1)
    return t, {k: v / n for k, v in tot.items()}


def step_bytes(nb: int, nv: int, contacts: float, word: int) -> float:
    """Compulsory HBM bytes of one whole step: every body's state and
    materials read, pos/theta/vel/omega written ((10 + 6) words), its local
    hull read (2 words per vertex), and the warm-start cache read and
    written (packed key + two FP64 lambdas = 24 B each way per contact)."""
    return nb * 16 * word + nv * 2 * word + contacts * 48


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` measured with
    rocprofv3 FETCH_SIZE and WRITE_SIZE counters (each scaled by the same
    counter on a copy of known size), from profiles/pmc_<kernel>_<workload>.md;
    None when no such measurement exists."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{workload}.md")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def roofline(scene, solver_name, steps, broadphase="auto", ppb=0, warm=engine.WARM_ZERO,
             model=None, threads=0):
    t, st = instrumented_run(scene, SOLVERS[solver_name], steps, broadphase, ppb, warm, model,
                             threads)
    nb = scene.n_bodies
    nv = sum(int(w.hull_off[-1]) for w in scene.worlds)
    word = 8 if solver_name == "ordered" else 4
    st["passes"] = scene.config.iterations + 1
    phases = [p for p in engine.PHASES if p not in ("reset",)]
    dom = max(phases, key=lambda p: t[p])
    ms = t[dom]
    by = kernel_bytes(dom, st, nb, nv, 8 if dom == "solve" else word)
    achieved = by / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    kname = f"solve_{solver_name}" if dom == "solve" else dom
    return {
        "kernel": dom,
        "kernels": PHASE_KERNELS.get((dom, solver_name), PHASE_KERNELS.get((dom, None), [])),
        "bound": "hbm",
        "achieved": round(achieved, 3),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 6),
        "traffic": measured_traffic(kname, scene.name),
        "algorithmic_bytes_per_launch": round(by),
        "avg_launch_ms": round(ms, 5),
    }, {**{p: round(t[p], 5) for p in phases}, "solve_profile_us": t["solve_profile_us"]}, st


def copy_calibration(nbytes=1 << 30, reps=3):
    """A plain device-to-device copy of a known size: what FETCH_SIZE and
    WRITE_SIZE read per dispatch for it gives their scale to bytes."""
    import torch
    x = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    x.fill_(1)
    torch.cuda.synchronize()
    for _ in range(reps):
        y = x.clone()
        del y
    torch.cuda.synchronize()
    print(f"copied {nbytes} bytes x{reps}", flush=True)


# ------------------------------------------------------------ CPU baseline

def cpu_baseline(scene, sample_steps, max_worlds=64):
    """The C++ FP64 oracle on this host: one world on one thread, or for a
    batch the first `max_worlds` worlds over the host's cores (at most 16)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    oracle.build()
    worlds = scene.worlds[:max_worlds]
    threads = 1 if len(worlds) == 1 else max(1, min(16, os.cpu_count() or 1, len(worlds)))
    engines = [oracle.OracleEngine(w, scene.config) for w in worlds]
    secs = oracle.run_engines(engines, sample_steps, threads=threads)
    for e in engines:
        e.close()
    dyn = sum(w.n_dynamic for w in worlds)
    what = (f"{scene.name} ({dyn} dynamic bodies)" if len(worlds) == 1 else
            f"worlds 0-{len(worlds) - 1} of {scene.name} ({dyn} dynamic bodies)")
    return {
        "value": round(dyn * sample_steps / secs, 1),
        "unit": "body-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {sample_steps} of {scene.config.steps} steps of {what}, "
                  f"C++ FP64 oracle, {threads} thread(s), {secs:.1f} s",
    }


BROADPHASES = {"auto": engine.BROADPHASE_AUTO, "scan": engine.BROADPHASE_SCAN,
               "grid": engine.BROADPHASE_GRID}


def autotune(scene, solver, ppb, steps=20, warm_steps=10):
    """Pick the schedule by measurement: broadphase (wavefront scan or
    uniform grid) x lanes per world in the solve (256 or 512, and 1024 for
    the coloured solver).  Each candidate steps the workload from its
    initial state (warm_steps untimed, then `steps` timed graph replays);
    the fastest wins."""
    times = {}
    lanes = (256, 512, 1024) if solver == engine.SOLVER_COLORED_F32 else (256, 512)
    for name in ("scan", "grid"):
        for threads in lanes:
            b = engine.make_batch(scene.worlds, scene.config, solver=solver,
                                  broadphase=BROADPHASES[name], max_pairs_per_body=ppb,
                                  solver_threads=threads)
            b.step_async(warm_steps, use_graph=True)
            b.sync()
            t0 = time.perf_counter()
            b.step_async(steps, use_graph=True)
            b.sync()
            times[(name, threads)] = (time.perf_counter() - t0) / steps * 1e3
            b.close()
    best = min(times, key=times.get)
    return best, {f"{k[0]}/{k[1]}": round(v, 4) for k, v in times.items()}


def run_mode(scene, solver, warm, steps, warmup, graph=True, model=None, broadphase="auto",
             ppb=0, threads=0):
    """Warm up on a separate instance, then time `steps` steps from the
    initial state; returns (seconds, run stats)."""
    def make():
        b = engine.make_batch(scene.worlds, scene.config, solver=solver, warm_start=warm,
                              broadphase=BROADPHASES[broadphase], max_pairs_per_body=ppb,
                              solver_threads=threads)
        if warm == engine.WARM_LEARNED:
            b.set_mlp(model)
        return b

    if warmup > 0:
        wb = make()
        wb.step_async(warmup, use_graph=graph)
        wb.sync()
        wb.close()
    b = make()
    b.sync()
    t0 = time.perf_counter()
    b.step_async(steps, use_graph=graph)
    b.sync()
    elapsed = time.perf_counter() - t0
    st = b.stats()
    b.close()
    return elapsed, st


def final_positions(scene, solver, warm, iterations, steps, model=None, broadphase="auto",
                    ppb=0, threads=0, worlds=None):
    """Positions (x, y) of every body of every world after `steps` steps, and
    the run's quality."""
    worlds = worlds if worlds is not None else scene.worlds
    cfg = scenes.SceneConfig(**{**scene.config.__dict__, "iterations": iterations})
    b = engine.make_batch(worlds, cfg, solver=solver, warm_start=warm,
                          broadphase=BROADPHASES[broadphase], max_pairs_per_body=ppb,
                          solver_threads=threads)
    if warm == engine.WARM_LEARNED:
        b.set_mlp(model)
    b.step_async(steps, use_graph=True)
    b.sync()
    st = b.stats()
    pos = [b.bodies(k)[:, :2].copy() for k in range(len(worlds))]
    b.close()
    return pos, quality(st)


def divergence(pa, pb, worlds):
    """Distance between two runs' final positions over the dynamic bodies."""
    d = np.concatenate([np.hypot(a[:, 0] - b[:, 0], a[:, 1] - b[:, 1])[(w.mass != 0) | (w.inertia != 0)]
                        for a, b, w in zip(pa, pb, worlds)])
    return {"max": round(float(d.max()), 6), "rms": round(float(np.sqrt((d ** 2).mean())), 6)}


def fp64_reference(scene, steps, runs, bp, ppb, model):
    """The FP64 drop-in solver (the reference's order, bit-exact to the
    oracle) at the workload's sweeps from the same initial state: its
    quality, and the divergence of each measured run's final positions from
    it.  `chaos_floor` is the divergence of the FP64 run itself when every
    dynamic body's initial x-velocity is moved by 1e-12: what any change of
    arithmetic leads to in these scenes."""
    base_it = scene.config.iterations
    ref, q_ref = final_positions(scene, engine.SOLVER_ORDERED_F64, engine.WARM_ZERO, base_it,
                                 steps, broadphase=bp, ppb=ppb)
    moved = []
    for w in scene.worlds:
        m = w.copy()
        dynm = (m.mass != 0) | (m.inertia != 0)
        m.vx = m.vx + np.where(dynm, 1e-12, 0.0)
        moved.append(m)
    pert, _ = final_positions(scene, engine.SOLVER_ORDERED_F64, engine.WARM_ZERO, base_it, steps,
                              broadphase=bp, ppb=ppb, worlds=moved)
    out = {"solver": "ordered", "iterations": base_it, "steps": steps, "quality": q_ref,
           "chaos_floor": divergence(pert, ref, scene.worlds), "runs": []}
    for warm, it, threads in runs:
        pos, q = final_positions(scene, engine.SOLVER_COLORED_F32, warm, it, steps, model=model,
                                 broadphase=bp, ppb=ppb, threads=threads)
        out["runs"].append({"warm_start": "learned" if warm == engine.WARM_LEARNED else "zero",
                            "iterations": it, "divergence": divergence(pos, ref, scene.worlds),
                            "mean_step_max_penetration": q["mean_step_max_penetration"],
                            "mean_step_normal_residual": q["mean_step_normal_residual"]})
    return out


def quality(st):
    return {"max_penetration": st["run_max_depth"],
            "mean_step_max_penetration": st["mean_max_depth"],
            "mean_step_normal_residual": st["mean_normal_residual"],
            "mean_step_friction_residual": st["mean_friction_residual"],
            "normal_residual": st["run_normal_residual"],
            "friction_residual": st["run_friction_residual"]}


def note(msg: str) -> None:
    """Progress of a long run on stderr (stdout carries the JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


QUALITY_TOL = 0.02


def matches(q, ref, tol=QUALITY_TOL):
    """Run means of the per-step maximum penetration and normal residual
    within `tol` of the reference run's."""
    return (q["mean_step_max_penetration"] <= ref["mean_step_max_penetration"] * (1 + tol) and
            q["mean_step_normal_residual"] <= ref["mean_step_normal_residual"] * (1 + tol))


# quality is compared over the headline world and these further seeds of
# the same generator (not the warm-start model's training seeds): a single
# chaotic run's means move by a few percent with any change of arithmetic
QUALITY_SEEDS = (1003, 1004)


def quality_over_seeds(scene, extra, st, warm, model, bp, ppb, threads, steps):
    """The run means of the headline world's run `st` and of untimed runs of
    the `extra` worlds in the same configuration, averaged."""
    rows = [quality(st)]
    for sc in extra:
        sc.config.iterations = scene.config.iterations
        _, st_x = run_mode(sc, engine.SOLVER_COLORED_F32, warm, steps, 0, graph=True,
                           model=model, broadphase=bp, ppb=ppb, threads=threads)
        rows.append(quality(st_x))
    keys = ("mean_step_max_penetration", "mean_step_normal_residual",
            "mean_step_friction_residual")
    out = {k: float(np.mean([r[k] for r in rows])) for k in keys}
    out["per_world"] = [{k: r[k] for k in keys[:2]} for r in rows]
    return out


def headline(scene, args, model, bp, ppb, dyn, threads):
    """Plain coloured solver at the workload's sweeps, then the learned
    warm start at decreasing sweeps while its quality (averaged over the
    headline world and QUALITY_SEEDS) matches."""
    base_it = scene.config.iterations
    graph = not args.no_graph
    extra = [scenes.polygon_pile(seed=sd) for sd in QUALITY_SEEDS] \
        if args.workload == "polygon_pile" else []
    el, st = run_mode(scene, engine.SOLVER_COLORED_F32, engine.WARM_ZERO, args.steps,
                      args.warmup, graph=graph, broadphase=bp, ppb=ppb, threads=threads)
    plain = {"warm_start": "zero", "iterations": base_it,
             "body_steps_per_s": round(dyn * args.steps / el, 1),
             "ms_per_step": round(1e3 * el / args.steps, 4), **quality(st),
             "quality_over_seeds": quality_over_seeds(scene, extra, st, engine.WARM_ZERO, None,
                                                      bp, ppb, threads, args.steps)}
    note(f"plain coloured at {base_it} sweeps: {plain['body_steps_per_s']:.0f} body-steps/s")
    tried = []
    chosen = None
    for it in range(base_it, 0, -1):
        scene.config.iterations = it
        el_l, st_l = run_mode(scene, engine.SOLVER_COLORED_F32, engine.WARM_LEARNED, args.steps,
                              args.warmup, graph=graph, model=model, broadphase=bp, ppb=ppb,
                              threads=threads)
        row = {"warm_start": "learned", "iterations": it,
               "body_steps_per_s": round(dyn * args.steps / el_l, 1),
               "ms_per_step": round(1e3 * el_l / args.steps, 4), **quality(st_l),
               "quality_over_seeds": quality_over_seeds(scene, extra, st_l, engine.WARM_LEARNED,
                                                        model, bp, ppb, threads, args.steps)}
        row["matches_plain"] = matches(row["quality_over_seeds"], plain["quality_over_seeds"])
        tried.append(row)
        note(f"learned at {it} sweeps: {row['body_steps_per_s']:.0f} body-steps/s, "
             f"matches plain: {row['matches_plain']}")
        if row["matches_plain"]:
            chosen = (it, el_l, st_l, row)
        elif it < base_it:
            break
    if chosen is None:  # even the full sweep count does not match: report it
        row = next(r for r in tried if r["iterations"] == base_it)
        chosen = (base_it, dyn * args.steps / row["body_steps_per_s"], None, row)
    scene.config.iterations = chosen[0]
    return plain, tried, chosen


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="polygon_pile")
    ap.add_argument("--mode", default="headline", choices=["headline", "single"],
                    help="headline: learned vs plain at fixed quality; single: one configuration")
    ap.add_argument("--solver", default="colored", choices=list(SOLVERS))
    ap.add_argument("--warm", default="zero", choices=["zero", "learned"])
    ap.add_argument("--iterations", type=int, default=0,
                    help="solver sweeps (default: the workload's)")
    ap.add_argument("--model", default=os.path.join(ROOT, "shapes_amd", "warmstart_mlp.npy"))
    ap.add_argument("--cpu-sample-steps", type=int, default=300)
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="steps of the instrumented replay (default: --steps)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--broadphase", default="tune", choices=["tune"] + list(BROADPHASES),
                    help="tune: pick broadphase and solver lanes by measurement")
    ap.add_argument("--solver-threads", type=int, default=0, choices=[0, 256, 512, 1024],
                    help="lanes per world in the solve when not tuned (0: by world size)")
    ap.add_argument("--pairs-per-body", type=int, default=0,
                    help="pair capacity per body (0: library default)")
    ap.add_argument("--worlds", type=int, default=0,
                    help="batch workloads: number of worlds (default: the config's)")
    ap.add_argument("--fp64-reference", action="store_true",
                    help="single mode, coloured solver: also run the FP64 drop-in solver and "
                         "report quality and trajectory divergence against it (headline mode "
                         "always does)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device three times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
    args = ap.parse_args()

    if args.copy_calibration:
        copy_calibration()
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build steps worlds on one GPU only; sharding worlds "
              "or slabs over several GPUs is not implemented (run with --gpus 1).",
              file=sys.stderr)
        sys.exit(2)

    t_start = time.perf_counter()
    scene = make_scene(args.workload, args.worlds)
    note(f"scene {scene.name}: {scene.n_bodies} bodies in {len(scene.worlds)} world(s), "
         f"built in {time.perf_counter() - t_start:.1f} s")
    if args.iterations:
        scene.config.iterations = args.iterations
    headline_mode = args.mode == "headline"
    solver_name = "colored" if headline_mode else args.solver
    solver = SOLVERS[solver_name]
    warm = engine.WARM_LEARNED if (headline_mode or args.warm == "learned") else engine.WARM_ZERO
    model = None
    if warm == engine.WARM_LEARNED:
        if not os.path.exists(args.model):
            raise SystemExit(f"bench.py: no model at {args.model}")
        model = np.load(args.model)
    dyn = scene.n_dynamic
    ppb = args.pairs_per_body
    tuning = None
    bp = args.broadphase
    threads = args.solver_threads
    if bp == "tune":
        # time each candidate in the middle of the run, where the contacts
        # of a settled scene set the cost, not on the first falling steps
        (bp, threads), tuning = autotune(scene, solver, ppb, steps=30,
                                         warm_steps=min(args.steps // 2, 300))
        note(f"autotuned schedule: {bp}, {threads} lanes {tuning}")

    extra = {}
    if headline_mode:
        plain, tried, chosen = headline(scene, args, model, bp, ppb, dyn, threads)
        it, elapsed, _, row = chosen
        q = {k: v for k, v in row.items()
             if k not in ("warm_start", "iterations", "body_steps_per_s", "ms_per_step",
                          "matches_plain", "quality_over_seeds")}
        extra = {"plain_gpu": plain, "learned_runs": tried,
                 "vs_plain_gpu": round(row["body_steps_per_s"] / plain["body_steps_per_s"], 4),
                 "quality_tolerance": QUALITY_TOL,
                 "quality_compared_over": ("the timed world and polygon piles with seeds "
                                           + ", ".join(map(str, QUALITY_SEEDS))
                                           if args.workload == "polygon_pile" else "the timed world"),
                 "quality_matched": bool(row.get("matches_plain"))}
        st_last = None
        scene.config.iterations = plain["iterations"]
        extra["fp64_reference"] = fp64_reference(
            scene, args.steps, [(engine.WARM_ZERO, plain["iterations"], threads),
                                (engine.WARM_LEARNED, it, threads)], bp, ppb, model)
        scene.config.iterations = it
        note("FP64 reference runs done")
    else:
        elapsed, st_last = run_mode(scene, solver, warm, args.steps, args.warmup,
                                    graph=not args.no_graph, model=model, broadphase=bp,
                                    ppb=ppb, threads=threads)
        q = quality(st_last)
        if args.fp64_reference and solver == engine.SOLVER_COLORED_F32:
            extra["fp64_reference"] = fp64_reference(
                scene, args.steps, [(warm, scene.config.iterations, threads)], bp, ppb, model)
            note("FP64 reference runs done")
    value = dyn * args.steps / elapsed
    note(f"timed run: {args.steps} steps in {elapsed:.2f} s")
    prof_steps = args.profile_steps or args.steps
    roof, phases, avg = roofline(scene, solver_name, prof_steps, bp, ppb, warm, model, threads)
    note(f"instrumented replay: {prof_steps} steps")
    nv = sum(int(w.hull_off[-1]) for w in scene.worlds)
    sb = step_bytes(scene.n_bodies, nv, avg["contacts"], DTYPES_BYTES[solver_name])
    step_gbs = sb / (elapsed / args.steps) / 1e9

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[solver_name],
        "data": "synthetic",
        "config": {
            "workload": scene.name,
            "worlds": len(scene.worlds),
            "bodies": scene.n_bodies,
            "dynamic_bodies": dyn,
            "iterations": scene.config.iterations,
            "dt": scene.config.dt,
            "solver": solver_name,
            "warm_start": "learned" if warm == engine.WARM_LEARNED else "zero",
            "mode": args.mode,
            "graph": not args.no_graph,
            "broadphase": bp,
            "solver_threads": threads,
            "autotune_ms_per_step": tuning,
        },
        "quality": q,
        **extra,
        "roofline": roof,
        "step_roofline": {"bound": "hbm", "algorithmic_bytes_per_step": round(sb),
                          "achieved": round(step_gbs, 3), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 6)},
        "phase_ms": phases,
        "avg_pairs": round(avg["pairs"], 1),
        "avg_contacts": round(avg["contacts"], 1),
    }
    if st_last is not None:
        out["last_step"] = {"pairs": st_last["pairs"], "contacts": st_last["contacts"],
                            "warm_hits": st_last["warm_hits"], "misses": st_last["misses"],
                            "learned_starts": st_last["learned_starts"],
                            "max_levels": st_last["max_levels"]}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, args.cpu_sample_steps)
        note("cpu baseline done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
