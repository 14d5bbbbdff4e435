"""The C ABI library loads and exports every symbol include/shapes_gpu.h
declares; argument validation needs no GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from shapes_amd import engine, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "shapes_gpu.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(shp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    L = engine.load_library()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(engine.EXPORTS)
    assert L.shp_version().startswith(b"shapes-mi355x")


def _desc(w):
    d = engine.WorldDescC()
    keep = []
    d.n_bodies = w.n
    for name in ("px", "py", "theta", "vx", "vy", "omega", "mass", "inertia",
                 "mu", "bounce"):
        a = np.ascontiguousarray(getattr(w, name))
        keep.append(a)
        setattr(d, name, a.ctypes.data_as(C.POINTER(C.c_double)))
    ho = np.ascontiguousarray(w.hull_off, dtype=np.int32)
    hx = np.ascontiguousarray(w.hull_x)
    hy = np.ascontiguousarray(w.hull_y)
    keep += [ho, hx, hy]
    d.hull_off = ho.ctypes.data_as(C.POINTER(C.c_int32))
    d.hull_x = hx.ctypes.data_as(C.POINTER(C.c_double))
    d.hull_y = hy.ctypes.data_as(C.POINTER(C.c_double))
    return d, keep


def test_create_rejects_bad_arguments():
    L = engine.load_library()
    w = scenes.pyramid(rows=2).worlds[0]
    d, keep = _desc(w)
    h = C.c_void_p()
    cfg = engine.ConfigC(1 / 60, 0.01, 0.02, 0, -9.8, 4, 0, 0, 0, 0)
    assert L.shp_batch_create(None, 1, C.byref(cfg), C.byref(h)) == 1
    bad = engine.ConfigC(0.0, 0.01, 0.02, 0, -9.8, 4, 0, 0, 0, 0)
    assert L.shp_batch_create(C.byref(d), 1, C.byref(bad), C.byref(h)) == 1
    assert b"dt" in L.shp_last_error()
    bad = engine.ConfigC(1 / 60, 0.01, 0.02, 0, -9.8, 4, 7, 0, 0, 0)
    assert L.shp_batch_create(C.byref(d), 1, C.byref(bad), C.byref(h)) == 1
    bad = engine.ConfigC(1 / 60, 0.01, 0.02, 0, -9.8, 4, 0, 0, 0, 5)
    assert L.shp_batch_create(C.byref(d), 1, C.byref(bad), C.byref(h)) == 1
    # a 9-gon is beyond the kernels' hull size
    b = scenes._Builder()
    b.add([(math_cos, math_sin) for (math_cos, math_sin) in
           [(np.cos(2 * np.pi * k / 9), np.sin(2 * np.pi * k / 9)) for k in range(9)]],
          (0, 0), mass=1, inertia=1)
    d9, keep9 = _desc(b.build())
    assert L.shp_batch_create(C.byref(d9), 1, C.byref(cfg), C.byref(h)) == 4
    # a 2-gon is invalid
    b = scenes._Builder()
    b.add([(0, 0), (1, 0)], (0, 0), mass=1, inertia=1)
    d2, keep2 = _desc(b.build())
    assert L.shp_batch_create(C.byref(d2), 1, C.byref(cfg), C.byref(h)) == 1
    # the coloured solver addresses a world's records with 32-bit offsets:
    # pair capacities from 2^28 are refused before any device work
    ppb = (1 << 28) // w.n + 1
    assert ppb * w.n + 1024 < (1 << 29)
    big = engine.ConfigC(1 / 60, 0.01, 0.02, 0, -9.8, 4, engine.SOLVER_COLORED_F32, 0, ppb, 0)
    assert L.shp_batch_create(C.byref(d), 1, C.byref(big), C.byref(h)) == 4
    assert b"coloured" in L.shp_last_error()


# ---- the launch plan and its audit (host only) ------------------------------

LDS_PER_CU = 160 * 1024


def _check_launches(rows):
    for r in rows:
        assert 1 <= r["workgroups"] <= 2**31 - 1, r
        assert r["workgroups"] * r["workgroup_size"] <= 2**32 - 1, r
        assert 64 <= r["workgroup_size"] <= 1024, r
        assert r["lds_bytes"] <= LDS_PER_CU, r


@pytest.fixture(scope="module")
def pile_batch():
    # 1024 piles of 2k-32k random polygons, 11.3 M bodies (about 20 s to build)
    return scenes.pile_batch()


@pytest.mark.parametrize("ppb", [4, 8, 16])
def test_plan_combined_at_scale(pile_batch, ppb):
    """Every launch of a step of the 1024-pile batch, coloured solver, stays
    within the hardware's limits, every 32-bit index in range, and the
    footprint fits one MI355X's 288 GB."""
    sc = pile_batch
    st, msg, p, rows = engine.plan(sc.worlds, sc.config, solver=engine.SOLVER_COLORED_F32,
                                   max_pairs_per_body=ppb)
    assert st == 0, msg
    assert p["bodies"] == sc.n_bodies == 11344603
    assert p["pair_capacity"] == ppb * p["bodies"] + 1024
    assert p["contact_capacity"] == 2 * p["pair_capacity"]
    assert p["broadphase"] == engine.BROADPHASE_GRID
    assert p["launches"] == len(rows) <= 64
    _check_launches(rows)
    names = [r["name"] for r in rows]
    assert names.count("k_colour") == 1 and names.count("k_sweep_colours") == 1
    assert "k_mlp_infer" not in names
    print(f"ppb {ppb}: {p['device_bytes'] / 1e9:.1f} GB, {len(rows)} launches, largest "
          f"{p['max_workitems']} work-items, {p['max_lds_bytes']} B LDS")
    assert p["device_bytes"] < 288e9
    # the 16-byte record offsets of the coloured sweeps fit 32 bits
    assert 16 * p["pair_capacity"] < 2**32


def test_plan_grid_lanes_at_scale(pile_batch):
    """Either grid schedule of the 1024-pile batch passes the audit.  The
    16-lane schedule launches 16 work-items per body for the count and the
    fill; the other counts with 4 work-items per body and fills with 1."""
    sc = pile_batch
    for bp, per_body in ((engine.BROADPHASE_GRID_1, (4, 1)), (engine.BROADPHASE_GRID_16, (16, 16))):
        st, msg, p, rows = engine.plan(sc.worlds, sc.config, solver=engine.SOLVER_COLORED_F32,
                                       max_pairs_per_body=4, broadphase=bp)
        assert st == 0, msg
        _check_launches(rows)
        g = [r for r in rows if r["name"].startswith("k_grid_pairs")]
        assert len(g) == 2
        for r, lanes in zip(g, per_body):
            assert lanes * p["bodies"] <= r["workgroups"] * r["workgroup_size"]
            assert r["workgroups"] * r["workgroup_size"] < lanes * p["bodies"] + 256


def test_plan_record_path_at_scale(pile_batch):
    """With 1024-lane colouring workgroups on the 1024-pile batch the contact
    kernel writes the manifold records after the colouring (no packing
    pass); with 256 lanes, or with fewer worlds than CUs, the colouring runs
    beside the contact kernel and a packing pass writes them."""
    sc = pile_batch
    for threads, worlds, packs in ((1024, sc.worlds, False), (512, sc.worlds, False),
                                   (256, sc.worlds, True), (1024, sc.worlds[:255], True)):
        st, msg, p, rows = engine.plan(worlds, sc.config, solver=engine.SOLVER_COLORED_F32,
                                       max_pairs_per_body=4, solver_threads=threads)
        assert st == 0, msg
        _check_launches(rows)
        names = [r["name"] for r in rows]
        assert ("k_pack_manifolds" in names) == packs, (threads, len(worlds))
        assert names.index("k_colour") < names.index("k_contacts")


def test_plan_learned_combined_at_scale(pile_batch):
    """The learned warm start on the 1024-pile batch: the inference launch
    and the miss buffers are planned within the limits."""
    sc = pile_batch
    st, msg, p, rows = engine.plan(sc.worlds, sc.config, solver=engine.SOLVER_COLORED_F32,
                                   warm_start=engine.WARM_LEARNED, max_pairs_per_body=4)
    assert st == 0, msg
    _check_launches(rows)
    infer = [r for r in rows if r["name"] == "k_mlp_infer"]
    assert len(infer) == 1 and infer[0]["workgroups"] <= 512
    assert p["miss_capacity"] == min(2 * p["pair_capacity"], 4 * p["bodies"] + 4096)
    assert p["device_bytes"] < 288e9


def test_plan_refuses_capacities_beyond_32_bit_offsets():
    w = scenes.pyramid(rows=2).worlds[0]
    cfg = scenes.SceneConfig()
    ppb = (1 << 29) // w.n + 1
    st, msg, _, _ = engine.plan([w], cfg, max_pairs_per_body=ppb)
    assert st == 4 and "2^29" in msg
    ppb = (1 << 28) // w.n + 1
    st, msg, _, _ = engine.plan([w], cfg, solver=engine.SOLVER_COLORED_F32, max_pairs_per_body=ppb)
    assert st == 4 and "coloured" in msg
    st, msg, p, rows = engine.plan([w], cfg, max_pairs_per_body=ppb)
    assert st == 0, msg
    _check_launches(rows)


def test_plan_scans_cover_their_counts():
    """One workgroup scans up to 20480 counts; beyond, three launches whose
    first and last cover every count in 8192-count workgroups."""
    for n_worlds, expect_parts in ((4, 0), (64, None)):
        sc = scenes.world_batch_fast(n_worlds=n_worlds)
        st, msg, p, rows = engine.plan(sc.worlds, sc.config)
        assert st == 0, msg
        scans = [r for r in rows if r["name"] == "scan_pairs"]
        n = p["pair_capacity"] + 1
        if n <= 20480:
            assert len(scans) == 1 and scans[0]["workgroups"] == 1
        else:
            assert len(scans) == 3
            assert scans[0]["workgroups"] == scans[2]["workgroups"] == -(-n // 8192)


def test_circle_descriptions_validated_on_the_host():
    """A circle must have an empty hull range and a finite radius >= 0
    (validated before any device work)."""
    sc = scenes.balls(dims=(2, 2))
    w = sc.worlds[0]
    st, msg, p, rows = engine.plan([w], sc.config)
    assert st == 0, msg
    assert p["vertices"] == 4 * 3  # the floor and two boxes; circles have none
    bad = w.copy()
    bad.radius[0] = 0.3  # the floor keeps its four vertices
    st, msg, _, _ = engine.plan([bad], sc.config)
    assert st == 1 and "empty" in msg
    bad = w.copy()
    bad.radius[1] = -1.0
    st, msg, _, _ = engine.plan([bad], sc.config)
    assert st == 1 and "radius" in msg
    bad.radius[1] = float("nan")
    st, msg, _, _ = engine.plan([bad], sc.config)
    assert st == 1


def test_capacity_adds_empty_slots_to_the_plan():
    """World.new capacity: slots beyond the bodies count in the plan (no
    vertices); a capacity below the body count is refused."""
    w = scenes.random_world(3, count=10)
    cfg = scenes.SceneConfig()
    st, msg, p0, _ = engine.plan([w], cfg)
    assert st == 0, msg
    w2 = w.copy()
    w2.capacity = w.n + 7
    st, msg, p, rows = engine.plan([w2], cfg)
    assert st == 0, msg
    assert p["bodies"] == w.n + 7 and p["vertices"] == p0["vertices"]
    assert p["max_world_bodies"] == w.n + 7
    _check_launches(rows)
    w2.capacity = w.n - 1
    st, msg, _, _ = engine.plan([w2], cfg)
    assert st == 1 and "capacity" in msg
    L = engine.load_library()
    for name in ("shp_batch_append", "shp_batch_delete", "shp_batch_filled",
                 "shp_batch_read_slots", "shp_batch_num_vertices"):
        assert hasattr(L, name)
