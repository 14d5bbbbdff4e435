"""Drop-in (ordered FP64) mode on the GPU: the hand-derived known answers of
the contact update, restitution, the longer pile run, overflow handling and
cache validation, all through the C ABI."""
import numpy as np
import pytest

import kat_scenes as K
from shapes_amd import engine, scenes
from test_gpu_parity import assert_bitwise, compare_state, run_parity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(K.CASES))
def test_contact_update_known_answers_gpu(orc, name):
    """The GPU's drop-in step gives the hand-derived answers (within 1e-14)
    and the oracle's bits."""
    case = K.CASES[name]
    w = K.world(**case["world"])
    cfg = K.config(case["iterations"])
    gpu = engine.make_batch([w], cfg)
    ref = orc.OracleEngine(w, cfg)
    if case["cache"] is not None:
        gpu.set_cache(0, np.array(K.KEYS, dtype=np.int32), np.array(case["cache"]))
        ref.set_cache(np.array(K.KEYS, dtype=np.int32), np.array(case["cache"]))
    gpu.step(1)
    ref.step(1)
    c = gpu.contacts()
    assert c["keys"].tolist() == K.KEYS
    want_lam = np.array([[float(x) for x in r] for r in case["lam"]])
    np.testing.assert_allclose(c["lam"], want_lam, rtol=1e-14, atol=1e-15)
    want_v = np.array([float(x) for x in case["v_b"]])
    np.testing.assert_allclose(gpu.bodies()[0, 3:6], want_v, rtol=1e-14, atol=1e-15)
    compare_state(gpu, ref, 1)
    gpu.close()


@pytest.mark.parametrize("seed", [11, 12])
def test_restitution_bitwise(orc, seed):
    """Bodies with bounce in (0, 1] (bounceB with e > 0 on every contact):
    pairs, keys, lagrangians and state bit for bit over 120 steps."""
    w = scenes.random_world(seed, count=60, spacing=1.25)
    rng = np.random.default_rng(seed)
    w.bounce = rng.uniform(0.2, 1.0, size=w.n)
    w.vy = w.vy - 3.0 * (w.mass != 0)  # falling onto the floor and each other
    cfg = scenes.SceneConfig(iterations=4)
    ref = run_parity(orc, w, cfg, steps=120, every=4)
    assert ref.stats()["contacts"] > 40


def test_polygon_pile_bitwise_full_run(orc):
    """The 8192-polygon world over all 600 steps of its configuration,
    compared every 50 steps (pairs, keys, geometry, lagrangians, state)."""
    sc = scenes.polygon_pile()
    run_parity(orc, sc.worlds[0], sc.config, steps=600, every=50)


def test_pair_overflow_reported_and_state_kept():
    """With too little pair capacity (one pair per body) the step reports
    SHP_E_OVERFLOW; graph-replayed steps after the overflow stay within the
    capacity (every pair slot below it holds a pair of the step) and the
    batch can still be read back."""
    rng = np.random.Generator(np.random.Philox(5))
    w = scenes.grid_world(rng, 3000, cols=60, spacing=0.55, y0=1.0, square_frac=0.5, vel=0.0)
    cfg = scenes.SceneConfig(iterations=2)
    for bp, solver in ((engine.BROADPHASE_SCAN, engine.SOLVER_ORDERED_F64),
                       (engine.BROADPHASE_GRID, engine.SOLVER_ORDERED_F64),
                       (engine.BROADPHASE_GRID, engine.SOLVER_COLORED_F32)):
        print(f"overflow: broadphase {bp}, solver {solver}", flush=True)
        b = engine.make_batch([w], cfg, max_pairs_per_body=1, broadphase=bp, solver=solver)
        b.step_async(5, use_graph=True)
        with pytest.raises(engine.ShapesError) as ei:
            b.sync()
        assert ei.value.status == 5
        s = b.stats()
        assert s["flags"] & 1
        assert s["pairs"] == w.n * 1 + 1024
        p = b.pairs(0)
        assert len(p) == w.n + 1024
        assert (p[:, 0] > p[:, 1]).all() and (p[:, 1] >= 0).all() and (p[:, 0] < w.n).all()
        assert np.isfinite(b.bodies(0)).all()
        b.close()


def test_set_cache_rejects_out_of_range_keys():
    w = scenes.pyramid(rows=3).worlds[0]
    b = engine.make_batch([w], scenes.SceneConfig())
    lam = np.zeros((1, 2))
    for bad in ([[w.n, 0, 0, 0]], [[1, w.n, 0, 0]], [[1, 0, 256, 0]], [[1, -1, 0, 0]]):
        with pytest.raises(engine.ShapesError) as ei:
            b.set_cache(0, np.array(bad, dtype=np.int32), lam)
        assert ei.value.status == 1
    b.set_cache(0, np.array([[w.n - 1, 0, 0, 0]], dtype=np.int32), lam)
    b.step(1)
    b.close()


@pytest.mark.parametrize("solver", ["ordered", "colored"])
def test_manifold_count_is_pairs_with_contacts(solver):
    """The step's manifold count (pairs with stored contacts) equals the
    number of distinct body pairs among the step's contact keys, on a
    batch of random worlds."""
    worlds = [scenes.random_world(40 + k, count=30 + 15 * k, spacing=1.25) for k in range(4)]
    sv = engine.SOLVER_ORDERED_F64 if solver == "ordered" else engine.SOLVER_COLORED_F32
    b = engine.make_batch(worlds, scenes.SceneConfig(iterations=4), solver=sv)
    for _ in range(3):
        b.step(20)
        s = b.stats()
        pairs = sum(len({(int(i), int(j)) for i, j, _, _ in b.contacts(k)["keys"]})
                    for k in range(len(worlds)))
        assert s["manifolds"] == pairs > 0
    b.close()


def test_pair_overflow_with_records_from_the_contact_kernel():
    """The coloured solver with 1024-lane colouring on 300 worlds (the
    contact kernel writes the manifold records) and too little pair
    capacity (one pair per body): the step reports SHP_E_OVERFLOW, the
    pairs stay within the capacity and every body stays finite over
    graph-replayed steps."""
    rng = np.random.Generator(np.random.Philox(6))
    worlds = [scenes.grid_world(rng, 40, cols=8, spacing=0.55, y0=1.0, square_frac=0.5, vel=0.0)
              for _ in range(300)]
    cfg = scenes.SceneConfig(iterations=2)
    st, msg, _, rows = engine.plan(worlds, cfg, solver=engine.SOLVER_COLORED_F32,
                                   max_pairs_per_body=1, solver_threads=1024)
    assert st == 0, msg
    assert "k_pack_manifolds" not in [r["name"] for r in rows]
    b = engine.make_batch(worlds, cfg, solver=engine.SOLVER_COLORED_F32, max_pairs_per_body=1,
                          solver_threads=1024)
    b.step_async(5, use_graph=True)
    with pytest.raises(engine.ShapesError) as ei:
        b.sync()
    assert ei.value.status == 5
    s = b.stats()
    n = sum(w.n for w in worlds)
    assert s["flags"] & 1 and s["pairs"] == n + 1024
    assert all(np.isfinite(b.bodies(k)).all() for k in range(len(worlds)))
    b.close()
