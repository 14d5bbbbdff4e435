"""The batched configurations at their full sizes on the GPU, through the C
ABI: every world of the batch is stepped, sampled worlds are compared with
the oracle, and size-independent properties are checked on all of them."""
import numpy as np
import pytest

from shapes_amd import engine, scenes
from test_gpu_parity import compare_state

pytestmark = pytest.mark.gpu


def _sample(scene, k):
    """k worlds spread over the batch's size ranks (smallest to largest)."""
    sizes = np.array([w.n for w in scene.worlds])
    order = np.argsort(sizes, kind="stable")
    return sorted(order[np.linspace(0, len(order) - 1, k).astype(int)].tolist())


@pytest.fixture(scope="module")
def world_batch():
    return scenes.world_batch_fast()  # 4096 worlds of 8-1024 bodies


def test_world_batch_full_run_ordered(orc, world_batch):
    """4096 worlds, 6 sweeps, all 300 steps in the drop-in solver: pairs,
    keys, geometry, lagrangians and state of 12 sampled worlds bit for bit
    at steps 100, 200 and 300."""
    sc = world_batch
    print(f"world batch: {sc.n_bodies} bodies, ordered, 300 steps", flush=True)
    gpu = engine.make_batch(sc.worlds, sc.config, max_pairs_per_body=4)
    sample = _sample(sc, 12)
    refs = {k: orc.OracleEngine(sc.worlds[k], sc.config) for k in sample}
    for step in (100, 200, 300):
        gpu.step_async(100, use_graph=True)
        gpu.sync()
        for k, r in refs.items():
            r.step(100)
            compare_state(gpu, r, step, world=k)
    gpu.close()


@pytest.mark.parametrize("lanes", [256, 64])
def test_world_batch_full_run_colored(orc, world_batch, lanes):
    """The same batch in the coloured FP32 solver (256 lanes per world, or
    one wavefront per world, several per workgroup) against the oracle's
    restatement of the coloured schedule on 12 sampled worlds, bit for bit
    at steps 100, 200, 300."""
    sc = world_batch
    print(f"world batch: {sc.n_bodies} bodies, coloured, {lanes} lanes, 300 steps", flush=True)
    gpu = engine.make_batch(sc.worlds, sc.config, solver=engine.SOLVER_COLORED_F32,
                            max_pairs_per_body=4, solver_threads=lanes)
    sample = _sample(sc, 12)
    refs = {k: orc.OracleEngine(sc.worlds[k], sc.config, solver=orc.SOLVER_COLORED, world_index=k)
            for k in sample}
    for step in (100, 200, 300):
        gpu.step_async(100, use_graph=True)
        gpu.sync()
        assert gpu.stats()["flags"] == 0
        for k, r in refs.items():
            r.step(100)
            gb, rb = gpu.bodies(k), r.bodies()
            assert np.array_equal(gb.view(np.uint64), rb.view(np.uint64)), f"world {k} step {step}"
            assert np.array_equal(gpu.contacts(k)["keys"], r.contacts()["keys"])
    s = gpu.stats()
    print("coloured world batch after 300 steps:", s)
    # run means of the per-step maxima (the run maxima are single impacts)
    assert s["mean_max_depth"] < 0.6
    gpu.close()


@pytest.fixture(scope="module")
def pile_batch():
    return scenes.pile_batch()  # 1024 piles of 2k-32k polygons, 11.3 M bodies


def test_pile_batch_full_size_colored(orc, pile_batch):
    """1024 worlds of 2k-32k bodies (worlds above ~13k bodies colour and
    keep their velocities in global memory), coloured solver, 4 sweeps,
    dt 1/120: 300 graph-replayed steps; three sampled worlds of up to 4000
    bodies bit for bit against the oracle's coloured restatement, no flag
    raised, every body finite and the run mean of the per-step maximum
    penetration bounded."""
    sc = pile_batch
    print(f"pile batch: {sc.n_bodies} bodies, coloured, 300 steps", flush=True)
    gpu = engine.make_batch(sc.worlds, sc.config, solver=engine.SOLVER_COLORED_F32,
                            max_pairs_per_body=4)
    small = [k for k in range(len(sc.worlds)) if sc.worlds[k].n <= 4000]
    sample = [small[0], small[len(small) // 2], small[-1]]
    refs = {k: orc.OracleEngine(sc.worlds[k], sc.config, solver=orc.SOLVER_COLORED, world_index=k)
            for k in sample}
    for step in (150, 300):
        gpu.step_async(150, use_graph=True)
        gpu.sync()
        for k, r in refs.items():
            r.step(150)
            gb, rb = gpu.bodies(k), r.bodies()
            assert np.array_equal(gb.view(np.uint64), rb.view(np.uint64)), f"world {k} step {step}"
    s = gpu.stats()
    print("pile batch after 300 steps:", s)
    assert s["flags"] == 0
    assert s["mean_max_depth"] < 0.5
    big = int(np.argmax([w.n for w in sc.worlds]))
    assert np.isfinite(gpu.bodies(big)).all()
    gpu.close()
