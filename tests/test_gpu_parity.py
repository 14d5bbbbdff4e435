"""Drop-in (ordered FP64) mode on the GPU against the CPU oracle.

Broadphase pairs, contact feature keys and orientation must match exactly;
contact geometry, lagrangians and body state must match bit for bit (the
1e-9 relative bound of the drop-in contract is met with zero error).  Every
call goes through the C ABI of libshapes_gpu.so.
"""
import numpy as np
import pytest

from shapes_amd import engine, scenes

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def assert_bitwise(name, got, want, step):
    got = np.asarray(got)
    want = np.asarray(want)
    assert got.shape == want.shape, f"{name} shape {got.shape} != {want.shape} at step {step}"
    if got.dtype.kind == "f":
        diff = np.nonzero(bits(got).ravel() != bits(want).ravel())[0]
    else:
        diff = np.nonzero(got.ravel() != want.ravel())[0]
    if len(diff):
        k = diff[0]
        raise AssertionError(f"{name} differs at step {step}, flat index {k}: "
                             f"gpu {got.ravel()[k]!r} oracle {want.ravel()[k]!r} "
                             f"({len(diff)} values differ)")


def compare_state(gpu, ref, step, world=0):
    flags = gpu.stats()["flags"]
    assert flags == 0, f"step flags {flags:#x} at step {step}"
    assert_bitwise("pairs", gpu.pairs(world), ref.pairs(), step)
    gc, rc = gpu.contacts(world), ref.contacts()
    for k in ("keys", "flip", "geom", "lam", "warm"):
        assert_bitwise(k, gc[k], rc[k], step)
    assert_bitwise("bodies", gpu.bodies(world), ref.bodies(), step)


def run_parity(orc, world, cfg, steps, every=1):
    print(f"parity: {world.n} bodies, {steps} steps", flush=True)
    gpu = engine.make_batch([world], cfg)
    ref = orc.OracleEngine(world, cfg)
    assert_bitwise("hulls", gpu.hulls(), ref.hulls(), 0)
    done = 0
    while done < steps:
        k = min(every, steps - done)
        gpu.step(k)
        ref.step(k)
        done += k
        compare_state(gpu, ref, done)
    assert_bitwise("hulls", gpu.hulls(), ref.hulls(), done)
    gpu.close()
    return ref


def test_transform_matches_oracle(orc):
    """sin/cos, the affine transform, sqrt and division of moveShapes on
    2000 spinning bodies far apart (no contacts), bit for bit."""
    rng = np.random.default_rng(0)
    b = scenes._Builder()
    for k in range(2000):
        verts = scenes.random_polygon(rng)
        b.add_dynamic(verts, ((k % 50) * 3.0, (k // 50) * 3.0 + 100.0),
                      theta=rng.uniform(-50, 50), omega=rng.uniform(-20, 20),
                      vel=(rng.uniform(-1, 1), rng.uniform(-1, 1)))
    w = b.build()
    cfg = scenes.SceneConfig(iterations=2)
    run_parity(orc, w, cfg, steps=20, every=5)


def test_pyramid_bitwise(orc):
    """Pyramid of 55 unit boxes, 4 sweeps, 300 steps, every step compared."""
    sc = scenes.pyramid()
    ref = run_parity(orc, sc.worlds[0], sc.config, steps=300)
    assert ref.stats()["contacts"] > 200


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_random_world_bitwise(orc, seed):
    w = scenes.random_world(seed, count=60, spacing=1.25)
    cfg = scenes.SceneConfig(iterations=4)
    run_parity(orc, w, cfg, steps=150)


def test_reference_bench_scene_bitwise(orc):
    """The reference's own benchmark scene, Stacks.makeScene (30,30) 0 with
    runWorld 0.01 ... 10 (shapes/bench/Main.hs:19-26): 901 bodies."""
    sc = scenes.reference_stacks()
    run_parity(orc, sc.worlds[0], sc.config, steps=10)


def test_batch_of_worlds_bitwise(orc):
    worlds = [scenes.random_world(100 + k, count=10 + 7 * k, spacing=1.3)
              for k in range(8)]
    cfg = scenes.SceneConfig(iterations=6)
    gpu = engine.make_batch(worlds, cfg)
    refs = [orc.OracleEngine(w, cfg) for w in worlds]
    for step in range(1, 81):
        gpu.step(1)
        for k, r in enumerate(refs):
            r.step(1)
            if step % 10 == 0:
                compare_state(gpu, r, step, world=k)
    gpu.close()


def test_polygon_pile_bitwise(orc):
    """The 8192-polygon world at full size while the first rows land."""
    sc = scenes.polygon_pile()
    run_parity(orc, sc.worlds[0], sc.config, steps=60, every=20)


def test_graph_replay_matches_eager():
    """A captured HIP graph replayed step after step gives the same bits as
    eager launches."""
    worlds = [scenes.random_world(7 + k, count=40, spacing=1.3) for k in range(4)]
    cfg = scenes.SceneConfig(iterations=4)
    a = engine.make_batch(worlds, cfg)
    b = engine.make_batch(worlds, cfg)
    for _ in range(6):
        a.step(10)
        b.step_async(10, use_graph=True)
        b.sync()
        for k in range(len(worlds)):
            assert_bitwise("bodies", b.bodies(k), a.bodies(k), 0)
            assert_bitwise("lam", b.contacts(k)["lam"], a.contacts(k)["lam"], 0)
    a.close()
    b.close()


def _run_grid(orc, worlds, cfg, steps, every):
    gpu = engine.make_batch(worlds, cfg, broadphase=engine.BROADPHASE_GRID)
    refs = [orc.OracleEngine(w, cfg) for w in worlds]
    done = 0
    while done < steps:
        gpu.step(every)
        for k, r in enumerate(refs):
            r.step(every)
            compare_state(gpu, r, done + every, world=k)
        done += every
    gpu.close()


def test_grid_broadphase_bitwise_pile(orc):
    """The uniform-grid broadphase gives the same pairs (set and order) and
    therefore the same bits as the scan, on the 8192-polygon world."""
    sc = scenes.polygon_pile()
    _run_grid(orc, sc.worlds, sc.config, steps=60, every=20)


def test_grid_broadphase_bitwise_worlds(orc):
    worlds = [scenes.random_world(500 + k, count=30 + 9 * k, spacing=1.25) for k in range(6)]
    worlds.append(scenes.pyramid().worlds[0])
    worlds.append(scenes.reference_stacks().worlds[0])
    _run_grid(orc, worlds, scenes.SceneConfig(iterations=3), steps=40, every=10)


@pytest.mark.parametrize("posts,boxes,steps", [(40, 50, 60), (240, 300, 12)])
def test_plank_chain_bitwise(orc, posts, boxes, steps):
    """A dynamic plank on static posts under loose boxes: the plank touches
    every contact, so its items form one chain through all passes (the
    longest wait the dataflow schedule can see).  The larger case has more
    contacts than the solver has lanes and a chain deeper than the depth
    sort takes, so it also runs the fallback lane order."""
    b = scenes._Builder()
    post = scenes.rectangle_vertices(0.6, 1.0)
    length = 1.5 * posts
    for k in range(posts):
        b.add(post, (-length / 2 + 0.75 + 1.5 * k, 0.5), mu=0.6)
    b.add_dynamic(scenes.rectangle_vertices(length, 0.5), (0.0, 1.245), mu=0.6)
    box = scenes.rectangle_vertices(0.9, 0.9)
    for k in range(boxes):
        b.add_dynamic(box, (-length / 2 + 0.6 + 1.2 * k * (length - 1.2) / (1.2 * boxes), 1.94),
                      mu=0.4)
    cfg = scenes.SceneConfig(iterations=8)
    ref = run_parity(orc, b.build(), cfg, steps=steps, every=3)
    assert ref.stats()["contacts"] >= 2 * (posts + boxes)


def test_large_world_global_counters_bitwise(orc):
    """24000 overlapping polygons in one world: more bodies than the
    per-body counters of LDS hold, so the ordered solver keeps them in
    global memory."""
    rng = np.random.Generator(np.random.Philox(77))
    w = scenes.grid_world(rng, 24000, cols=200, spacing=1.05, y0=1.0, square_frac=0.5, vel=0.3)
    cfg = scenes.SceneConfig(iterations=3)
    ref = run_parity(orc, w, cfg, steps=4, every=2)
    assert ref.stats()["contacts"] > 20000


def test_grid_broadphase_bitwise_large_batch(orc):
    """A batch of ~150k bodies takes the one-lane-per-body grid kernel;
    pairs and state of a sample of its worlds match the oracle bit for bit."""
    sc = scenes.world_batch_fast(n_worlds=720)
    assert sc.n_bodies > 131072
    gpu = engine.make_batch(sc.worlds, sc.config, broadphase=engine.BROADPHASE_GRID)
    sample = list(range(0, 720, 37)) + [719]
    refs = {k: orc.OracleEngine(sc.worlds[k], sc.config) for k in sample}
    for step in (5, 10):
        gpu.step(5)
        for k, r in refs.items():
            r.step(5)
            compare_state(gpu, r, step, world=k)
    gpu.close()
