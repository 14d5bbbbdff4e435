"""World mutation on the GPU (World.append / World.delete over the
EmptiesVector free list, shapes/src/Physics/World.hs:77-91,
shapes/src/Utils/EmptiesVector.hs:27-92) against the oracle doing the same
operations, bit for bit, through the C ABI."""
import math

import numpy as np
import pytest

from shapes_amd import engine, scenes
from test_gpu_parity import assert_bitwise, compare_state

pytestmark = pytest.mark.gpu


def ngon(k, r):
    return [(r * math.cos(2 * math.pi * q / k + 0.1), r * math.sin(2 * math.pi * q / k + 0.1))
            for q in range(k)]


def append_both(gpu, refs, world, **body):
    s = gpu.append(world, **body)
    assert refs[world].append(**body) == s
    return s


def delete_both(gpu, refs, world, slot):
    gpu.delete(world, slot)
    assert refs[world].delete(slot) == 0


def check(gpu, refs, step, worlds, stepped=True):
    """After a step: everything; right after a mutation (stepped False):
    the state, hulls and slot table (a re-laid batch reads back no pairs
    or contacts until its next step)."""
    for k in worlds:
        if stepped:
            compare_state(gpu, refs[k], step, world=k)
        else:
            assert_bitwise("bodies", gpu.bodies(k), refs[k].bodies(), step)
        assert_bitwise("hulls", gpu.hulls(k), refs[k].hulls(), step)
        assert gpu.slots(k).tolist() == refs[k].slots().tolist()
        assert gpu.filled(k) == refs[k].filled()


def vertex_count(w, i):
    return int(w.hull_off[i + 1] - w.hull_off[i])


@pytest.mark.parametrize("broadphase", [engine.BROADPHASE_GRID, engine.BROADPHASE_SCAN])
def test_mutation_dropin_bitwise(orc, broadphase):
    """Circles and polygons with six spare slots: step, delete three
    bodies, step, append five (into the emptied slots in LIFO order, then
    the never-filled ones: in place when the vertex count matches, else by
    laying the arrays out again), step, with every step's pairs, contacts,
    lagrangians, state, hulls and slot table equal to the oracle's."""
    w = scenes.circle_world(7, count=50)
    w.capacity = w.n + 6
    cfg = scenes.SceneConfig(iterations=4)
    print(f"mutation: {w.n} bodies, capacity {w.capacity}, broadphase {broadphase}", flush=True)
    gpu = engine.make_batch([w], cfg, broadphase=broadphase)
    refs = [orc.OracleEngine(w, cfg)]
    step = 0
    for _ in range(15):
        gpu.step(1)
        refs[0].step(1)
        step += 1
        check(gpu, refs, step, [0])
    gone = [5, 12, 30]
    for s in gone:
        delete_both(gpu, refs, 0, s)
    for _ in range(10):
        gpu.step(1)
        refs[0].step(1)
        step += 1
        check(gpu, refs, step, [0])
    # slot 30 first (LIFO): a small body with its old vertex count (in place)
    k30 = vertex_count(w, 30)
    same = dict(hull=ngon(k30, 0.2)) if k30 else dict(radius=0.2)
    assert append_both(gpu, refs, 0, px=-2.0, py=9.0, vx=0.3, mass=0.5, inertia=0.05,
                       **same) == 30
    # slot 12: a different vertex count (laid out again)
    k12 = vertex_count(w, 12)
    other = dict(hull=ngon(7 if k12 != 7 else 5, 0.3))
    assert append_both(gpu, refs, 0, px=0.0, py=10.0, omega=1.0, mass=0.8, inertia=0.1,
                       **other) == 12
    # slot 5, then the never-filled slots: a circle (no vertices: in place
    # into a never-filled slot), a static polygon
    assert append_both(gpu, refs, 0, px=2.0, py=9.5, radius=0.25, mass=0.4,
                       inertia=0.0125) == 5
    assert append_both(gpu, refs, 0, px=3.5, py=8.0, radius=0.3, mass=0.6,
                       inertia=0.027, bounce=0.4) == w.n
    assert append_both(gpu, refs, 0, px=-4.0, py=2.0, mass=0.0, inertia=0.0,
                       hull=ngon(4, 0.5)) == w.n + 1
    check(gpu, refs, step, [0], stepped=False)
    for _ in range(25):
        gpu.step(1)
        refs[0].step(1)
        step += 1
        check(gpu, refs, step, [0])
    gpu.close()


def test_mutation_coloured_batch_with_graph(orc):
    """Three worlds, coloured solver, graph replay: world 1 loses two
    bodies and gains three (one laying the arrays out again, which
    re-captures the graph); every world matches the oracle's restatement of
    the coloured schedule bit for bit, the unmutated ones included."""
    worlds = [scenes.random_world(30 + k, count=36) for k in range(3)]
    for w in worlds:
        w.capacity = w.n + 3
    cfg = scenes.SceneConfig(iterations=6)
    gpu = engine.make_batch(worlds, cfg, solver=engine.SOLVER_COLORED_F32)
    refs = [orc.OracleEngine(w, cfg, solver=orc.SOLVER_COLORED, world_index=k)
            for k, w in enumerate(worlds)]
    for r in refs:
        r.step(20)
    gpu.step_async(20, use_graph=True)
    gpu.sync()
    check(gpu, refs, 20, range(3))
    delete_both(gpu, refs, 1, 9)
    delete_both(gpu, refs, 1, 21)
    gpu.step_async(10, use_graph=True)
    for r in refs:
        r.step(10)
    gpu.sync()
    check(gpu, refs, 30, range(3))
    assert append_both(gpu, refs, 1, px=0.5, py=12.0, mass=1.0, inertia=0.2,
                       hull=ngon(8, 0.4)) == 21
    assert append_both(gpu, refs, 1, px=-1.5, py=12.0, mass=1.0, inertia=0.2,
                       hull=ngon(3, 0.4)) == 9
    assert append_both(gpu, refs, 1, px=2.5, py=13.0, radius=0.3, mass=0.5,
                       inertia=0.02) == worlds[1].n
    gpu.step_async(30, use_graph=True)
    for r in refs:
        r.step(30)
    gpu.sync()
    check(gpu, refs, 60, range(3))
    gpu.close()


def test_mutation_errors():
    """safeDelete of an empty slot and append into a full world are
    refused with SHP_E_INVAL, leaving the slot table as it was."""
    w = scenes.random_world(3, count=10)
    w.capacity = w.n + 1
    gpu = engine.make_batch([w], scenes.SceneConfig())
    assert gpu.filled() == w.n
    assert gpu.append(0, px=0.0, py=20.0, hull=ngon(4, 0.5)) == w.n
    with pytest.raises(engine.ShapesError) as e:
        gpu.append(0, px=0.0, py=22.0, hull=ngon(4, 0.5))
    assert e.value.status == 1
    gpu.delete(0, 4)
    with pytest.raises(engine.ShapesError) as e:
        gpu.delete(0, 4)
    assert e.value.status == 1
    with pytest.raises(engine.ShapesError):
        gpu.delete(0, w.capacity)
    with pytest.raises(engine.ShapesError):
        gpu.append(0, px=0.0, py=22.0, hull=ngon(9, 0.5))
    assert gpu.filled() == w.n
    assert gpu.slots().tolist() == [k != 4 for k in range(w.capacity)]
    gpu.step(5)
    gpu.close()


def test_relayout_keeps_learned_start_log_and_cache(orc):
    """Laying the arrays out again keeps the learned model, the training
    log (its count goes on growing) and the warm-start caches (the next
    step reuses cached lagrangians)."""
    sc = scenes.polygon_pile(seed=9, count=512)
    w = sc.worlds[0]
    w.capacity = w.n + 2
    gpu = engine.make_batch([w], sc.config, solver=engine.SOLVER_COLORED_F32,
                            warm_start=engine.WARM_LEARNED)
    import os
    model = np.load(os.path.join(os.path.dirname(engine.__file__), "warmstart_mlp.npy"))
    gpu.set_mlp(model)
    gpu.set_logging(1 << 20)
    gpu.step(60)
    n0 = len(gpu.read_log()[0])
    assert n0 > 0
    assert gpu.append(0, px=0.0, py=60.0, mass=1.0, inertia=0.2, hull=ngon(5, 0.4)) == w.n
    gpu.step(1)
    st = gpu.stats()
    assert st["flags"] == 0 and st["warm_hits"] > 0
    gpu.step(20)
    assert len(gpu.read_log()[0]) > n0
    assert gpu.stats()["learned_starts"] >= 0
    gpu.close()
