"""Circle shapes on the GPU (Contact.hs:22-40: circle-circle, circle-hull
through GJK, hull-circle), drop-in and coloured, against the oracle through
the C ABI."""
import numpy as np
import pytest

from shapes_amd import engine, scenes
from test_gpu_parity import compare_state, run_parity

pytestmark = pytest.mark.gpu


def test_circle_fixture_contacts_gpu(orc):
    """The reference's GJK fixture box (makeOptBox 0 0 1.9 1.9, static)
    under a circle of radius 0.5 centred at (0, 1): one contact keyed
    (0, 1) on the box's edge (1, 0) at (0, 0.95), depth 0.45, the same bits
    as the oracle."""
    b = scenes._Builder()
    b.add(scenes.rectangle_vertices(1.9, 1.9), (0.0, 0.0), mu=0.5)
    b.add_circle(0.5, (0.0, 1.0), mass=1.0, inertia=0.125, mu=0.5)
    w = b.build()
    cfg = scenes.SceneConfig(dt=1.0 / 64.0, gy=0.0, iterations=1)
    gpu = engine.make_batch([w], cfg)
    ref = orc.OracleEngine(w, cfg)
    gpu.step(1)
    ref.step(1)
    c = gpu.contacts()
    assert c["keys"].tolist() == [[1, 0, 0, 1]] and c["flip"].tolist() == [0]
    assert c["geom"][0, 2:4].tolist() == [0.0, 0.95]
    assert abs(c["geom"][0, 4] - 0.45) < 1e-15
    compare_state(gpu, ref, 1)
    gpu.close()


def test_balls_scene_bitwise(orc):
    """The reference's circle scene Balls.makeScene (10,10) 0.5 0
    (Balls.hs:68-71): 50 circles (bounce 0.5) and 50 boxes in alternating
    columns on the 18x1 floor, 2 sweeps, 300 steps, every tenth step
    compared bit for bit (pairs, keys, geometry, lagrangians, state)."""
    sc = scenes.balls()
    ref = run_parity(orc, sc.worlds[0], sc.config, steps=300, every=10)
    keys = ref.contacts()["keys"]
    rad = sc.worlds[0].radius
    assert (rad[keys[:, 0]] > 0).any() and (rad[keys[:, 1]] > 0).any()


@pytest.mark.parametrize("seed", [31, 32, 33])
def test_circle_world_bitwise(orc, seed):
    """Circles of random radii and bounce among random polygons, 4 sweeps,
    150 steps, bit for bit every 5 steps."""
    w = scenes.circle_world(seed)
    run_parity(orc, w, scenes.SceneConfig(iterations=4), steps=150, every=5)


def test_circles_grid_broadphase_and_batch(orc):
    """Circle worlds in a batch with the grid broadphase (a circle's cell
    extent is its diameter) give the oracle's bits."""
    worlds = [scenes.circle_world(40 + k, count=30 + 10 * k) for k in range(4)]
    worlds.append(scenes.balls(dims=(6, 8)).worlds[0])
    cfg = scenes.SceneConfig(iterations=3)
    gpu = engine.make_batch(worlds, cfg, broadphase=engine.BROADPHASE_GRID)
    refs = [orc.OracleEngine(w, cfg) for w in worlds]
    for step in range(10, 61, 10):
        gpu.step(10)
        for k, r in enumerate(refs):
            r.step(10)
            compare_state(gpu, r, step, world=k)
    gpu.close()


def test_circles_coloured_matches_restatement(orc):
    """The coloured FP32 solver on the circle scene against the oracle's
    restatement of the coloured schedule, bit for bit."""
    sc = scenes.balls()
    w = sc.worlds[0]
    gpu = engine.make_batch([w], sc.config, solver=engine.SOLVER_COLORED_F32)
    ref = orc.OracleEngine(w, sc.config, solver=orc.SOLVER_COLORED)
    for step in range(20, 201, 20):
        gpu.step(20)
        ref.step(20)
        gb, rb = gpu.bodies(), ref.bodies()
        assert np.array_equal(gb.view(np.uint64), rb.view(np.uint64)), f"step {step}"
    gpu.close()
