"""Graph-coloured FP32 solver: deterministic, same contact detection as the
reference on the first step, penetration and residual bounded, and close
to the FP64 drop-in trajectory over short horizons."""
import numpy as np
import pytest

from shapes_amd import engine, scenes

pytestmark = pytest.mark.gpu


def test_colored_is_deterministic():
    w = scenes.random_world(21, count=80, spacing=1.3)
    cfg = scenes.SceneConfig(iterations=6)
    runs = []
    for _ in range(2):
        b = engine.make_batch([w], cfg, solver=engine.SOLVER_COLORED_F32)
        b.step(100)
        runs.append(b.bodies())
        b.close()
    assert np.array_equal(runs[0].view(np.uint64), runs[1].view(np.uint64))


def test_colored_first_step_contacts_match_oracle(orc):
    sc = scenes.pyramid()
    w = sc.worlds[0]
    b = engine.make_batch([w], sc.config, solver=engine.SOLVER_COLORED_F32)
    r = orc.OracleEngine(w, sc.config)
    b.step(1)
    r.step(1)
    assert np.array_equal(b.pairs(), r.pairs())
    assert np.array_equal(b.contacts()["keys"], r.contacts()["keys"])
    b.close()


def test_colored_pyramid_stays_up(orc):
    sc = scenes.pyramid()
    w = sc.worlds[0]
    b = engine.make_batch([w], sc.config, solver=engine.SOLVER_COLORED_F32)
    r = orc.OracleEngine(w, sc.config)
    b.step(300)
    r.step(300)
    s = b.stats()
    print("coloured pyramid:", s)
    assert s["run_max_depth"] < 0.05
    assert s["max_levels"] < 64
    # the pyramid stands: every box within 3 cm of the FP64 reference after
    # 300 steps (the coloured schedule on pyramid seeds 1-8: 0.9-2.2 cm)
    d = np.abs(b.bodies()[:, :2] - r.bodies()[:, :2]).max()
    print("max position divergence vs FP64 reference:", d)
    assert d < 0.03
    b.close()


def test_ordered_and_colored_quality_on_batch():
    worlds = [scenes.random_world(300 + k, count=50, spacing=1.3) for k in range(16)]
    cfg = scenes.SceneConfig(iterations=8)
    out = {}
    for name, solver in (("ordered", engine.SOLVER_ORDERED_F64),
                         ("colored", engine.SOLVER_COLORED_F32)):
        b = engine.make_batch(worlds, cfg, solver=solver)
        b.step(200)
        out[name] = b.stats()
        b.close()
    print(out)
    assert out["colored"]["run_max_depth"] < 2.0 * out["ordered"]["run_max_depth"] + 0.02


def _colored_vs_restatement(orc, worlds, cfg, steps, every, threads=0):
    """GPU coloured solver against the oracle's restatement of the same
    schedule (Jones-Plassmann colours, FP32 rows and velocities, FP64
    accumulators, colour order): every bit of state must agree."""
    gpu = engine.make_batch(worlds, cfg, solver=engine.SOLVER_COLORED_F32,
                            solver_threads=threads)
    refs = [orc.OracleEngine(w, cfg, solver=orc.SOLVER_COLORED, world_index=k)
            for k, w in enumerate(worlds)]
    done = 0
    while done < steps:
        gpu.step(every)
        done += every
        assert gpu.stats()["flags"] == 0
        for k, r in enumerate(refs):
            r.step(every)
            gb, rb = gpu.bodies(k), r.bodies()
            diff = np.nonzero(gb.view(np.uint64).ravel() != rb.view(np.uint64).ravel())[0]
            assert len(diff) == 0, (f"world {k} step {done}: {len(diff)} values differ, first "
                                    f"{gb.ravel()[diff[0]]!r} vs {rb.ravel()[diff[0]]!r}")
            gl, rl = gpu.contacts(k)["lam"], r.contacts()["lam"]
            assert np.array_equal(gl.view(np.uint64), rl.view(np.uint64)), f"lambdas, world {k}"
    gpu.close()


def test_colored_matches_restatement_pyramid(orc):
    sc = scenes.pyramid()
    _colored_vs_restatement(orc, sc.worlds, sc.config, steps=300, every=10)


def test_colored_matches_restatement_batch(orc):
    worlds = [scenes.random_world(700 + k, count=20 + 11 * k, spacing=1.25) for k in range(6)]
    worlds.append(scenes.reference_stacks().worlds[0])
    _colored_vs_restatement(orc, worlds, scenes.SceneConfig(iterations=6), steps=60, every=10)


@pytest.mark.parametrize("threads", [256, 512, 1024])
def test_colored_matches_restatement_polygon_pile(orc, threads):
    """The 8192-polygon world (LDS velocities), with 4, 8 and 16 waves:
    colours wider than the lanes, so lanes take several manifolds of a
    colour."""
    sc = scenes.polygon_pile()
    _colored_vs_restatement(orc, sc.worlds, sc.config, steps=40, every=2, threads=threads)


def test_colored_matches_restatement_single_colour(orc):
    """Three boxes resting on one floor: every manifold shares only the
    static floor, so the world has one colour and each pass solves the same
    manifolds again."""
    b = scenes._Builder()
    b.add(scenes.rectangle_vertices(20.0, 1.0), (0.0, -0.5), mu=0.5)
    for x in (-5.0, 0.0, 5.0):
        b.add_dynamic(scenes.rectangle_vertices(1.0, 1.0), (x, 0.49), mu=0.5)
    _colored_vs_restatement(orc, [b.build()], scenes.SceneConfig(iterations=4), steps=60, every=5)


@pytest.mark.parametrize("n", [24000, 30000, 45000])
def test_colored_matches_restatement_global_path(orc, n):
    """Worlds too large for LDS, velocities in global memory: coloured with
    the bodies' priorities and colour masks in LDS (24000), the priorities
    only (30000), or entirely in global memory (45000)."""
    rng = np.random.Generator(np.random.Philox(78))
    w = scenes.grid_world(rng, n, cols=200, spacing=1.05, y0=1.0, square_frac=0.5, vel=0.3)
    print(f"global-path colouring: {n} bodies", flush=True)
    _colored_vs_restatement(orc, [w], scenes.SceneConfig(iterations=3), steps=4, every=2)


def test_wavefront_sweeps_mixed_sizes_bitwise(orc):
    """One wavefront per world for the worlds of at most 1024 bodies and a
    256-lane workgroup for the larger ones, in one batch: every world equals
    the oracle's restatement of the coloured schedule bit for bit."""
    worlds = [scenes.random_world(70 + k, count=n, cols=max(8, int(n ** 0.5)))
              for k, n in enumerate((1400, 300, 40, 1100, 900, 5))]
    cfg = scenes.SceneConfig(iterations=6)
    st, msg, p, rows = engine.plan(worlds, cfg, solver=engine.SOLVER_COLORED_F32,
                                   solver_threads=64)
    assert st == 0, msg
    names = {r["name"]: r for r in rows}
    assert names["k_sweep_colours"]["workgroups"] == 2  # the two worlds above 1024 bodies
    assert names["k_sweep_waves"]["workgroups"] == 1    # four small worlds, one workgroup
    gpu = engine.make_batch(worlds, cfg, solver=engine.SOLVER_COLORED_F32, solver_threads=64)
    refs = [orc.OracleEngine(w, cfg, solver=orc.SOLVER_COLORED, world_index=k)
            for k, w in enumerate(worlds)]
    for step in range(4):
        gpu.step_async(25, use_graph=True)
        gpu.sync()
        assert gpu.stats()["flags"] == 0
        for k, r in enumerate(refs):
            r.step(25)
            assert np.array_equal(gpu.bodies(k).view(np.uint64), r.bodies().view(np.uint64)), \
                f"world {k} step {25 * (step + 1)}"
            assert np.array_equal(gpu.contacts(k)["lam"].view(np.uint64),
                                  r.contacts()["lam"].view(np.uint64))
    gpu.close()


@pytest.mark.parametrize("warm", ["zero", "learned"])
def test_record_paths_agree(warm):
    """The coloured solver's manifold records written by the contact kernel
    after the colouring (1024-lane colouring workgroups, at least 256
    worlds) and by the packing pass after the colouring ran beside the
    contact kernel (256 lanes): the same bodies bit for bit over 40 steps,
    with the zero and with the learned warm start (whose predictions the
    first path patches into records already written)."""
    from shapes_amd import warmstart
    worlds = [scenes.random_world(900 + k, count=30 + 10 * (k % 7), spacing=1.25)
              for k in range(300)]
    cfg = scenes.SceneConfig(iterations=4)
    ws = engine.WARM_LEARNED if warm == "learned" else engine.WARM_ZERO
    out = []
    for threads, packs in ((1024, False), (256, True)):
        st, msg, _, rows = engine.plan(worlds, cfg, solver=engine.SOLVER_COLORED_F32,
                                       warm_start=ws, solver_threads=threads)
        assert st == 0, msg
        assert ("k_pack_manifolds" in [r["name"] for r in rows]) == packs
        print(f"record path test: {warm} start, {threads} lanes", flush=True)
        b = engine.make_batch(worlds, cfg, solver=engine.SOLVER_COLORED_F32, warm_start=ws,
                              solver_threads=threads)
        if ws == engine.WARM_LEARNED:
            b.set_mlp(np.load(warmstart.DEFAULT_MODEL))
        learned = 0
        for _ in range(40):
            b.step(1)
            learned += b.stats()["learned_starts"]
        assert b.stats()["flags"] == 0
        assert (learned > 0) == (ws == engine.WARM_LEARNED)
        out.append(np.concatenate([b.bodies(k) for k in range(len(worlds))]))
        b.close()
    assert np.array_equal(out[0].view(np.uint64), out[1].view(np.uint64))
