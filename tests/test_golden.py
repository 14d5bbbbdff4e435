"""Golden fixtures (tests/golden/, written by make_golden.py from the
oracle) in the dump format of shp_batch_dump: the oracle reproduces them,
they parse, and they hold the invariants of the quantities they record."""
import json
import os

import numpy as np

import golden_util
from shapes_amd import wire

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _c1():
    with open(os.path.join(HERE, "c1_pyramid.shpd"), "rb") as f:
        return f.read()


def test_oracle_reproduces_c1_golden(orc):
    assert golden_util.c1_oracle_dumps(orc) == _c1()


def test_oracle_reproduces_c2_golden_first_checkpoints(orc):
    with open(os.path.join(HERE, "c2_polygon_pile.json")) as f:
        want = json.load(f)
    got = golden_util.c2_oracle_digests(orc, upto=200)
    assert set(got) == {"1", "50", "100", "200"}
    for k, v in got.items():
        assert v == want[k], k


def _desc(keys):
    """Keys as tuples in strictly descending lexicographic order."""
    t = [tuple(int(x) for x in k) for k in keys]
    return all(a > b for a, b in zip(t, t[1:]))


def test_c1_golden_invariants():
    """Every dump parses and writes back to the same bytes; slots ascend,
    pairs and contact keys descend (i > j), the cache after a step is the
    step's contacts with their lagrangians; the probe's contacts start from
    the cached lagrangians; each AABB bounds its box's corners placed from
    the recorded state (to rounding)."""
    data = _c1()
    dumps = wire.read(data)
    assert [d["steps"] for d in dumps] == golden_util.C1_CHECKPOINTS + [300]
    pos = 0
    box = np.array([[-0.5, -0.5], [0.5, -0.5], [0.5, 0.5], [-0.5, 0.5]])
    for n, d in enumerate(dumps):
        b = wire.write(d["world"], d["steps"],
                       bodies=(d["BODY"]["slot"], np.stack([d["BODY"][k] for k in
                               ("px", "py", "theta", "vx", "vy", "omega")], 1)),
                       aabbs=(d["AABB"]["slot"], np.stack([d["AABB"][k] for k in
                              ("minx", "maxx", "miny", "maxy")], 1)),
                       pairs=np.stack([d["PAIR"]["i"], d["PAIR"]["j"]], 1),
                       contacts={"keys": np.stack([d["CONT"][k] for k in ("i", "j", "fi", "fj")], 1),
                                 "flip": d["CONT"]["flip"], "warm": d["CONT"]["warm"],
                                 "geom": np.stack([d["CONT"][k] for k in
                                                   ("nx", "ny", "px", "py", "depth")], 1),
                                 "lam": np.stack([d["CONT"]["ln"], d["CONT"]["lt"]], 1)},
                       cache=(np.stack([d["CACH"][k] for k in ("i", "j", "fi", "fj")], 1),
                              np.stack([d["CACH"]["ln"], d["CACH"]["lt"]], 1)))
        assert data[pos:pos + len(b)] == b
        pos += len(b)
        assert d["BODY"]["slot"].tolist() == list(range(56)) == d["AABB"]["slot"].tolist()
        p = d["PAIR"]
        assert (p["i"] > p["j"]).all() and _desc(np.stack([p["i"], p["j"]], 1))
        ck = np.stack([d["CONT"][k] for k in ("i", "j", "fi", "fj")], 1)
        hk = np.stack([d["CACH"][k] for k in ("i", "j", "fi", "fj")], 1)
        assert _desc(ck) and _desc(hk)
        if n < len(dumps) - 1:  # after a step: the cache is the step's solution
            assert np.array_equal(ck, hk)
            assert np.array_equal(d["CONT"]["ln"], d["CACH"]["ln"])
            assert np.array_equal(d["CONT"]["lt"], d["CACH"]["lt"])
            assert (d["CONT"]["ln"] >= 0).all()
        else:  # the probe: cached lagrangians where the key is cached, else 0
            cache = {tuple(k): (a, b) for k, a, b in zip(hk.tolist(), d["CACH"]["ln"],
                                                        d["CACH"]["lt"])}
            for k, w, a, b in zip(ck.tolist(), d["CONT"]["warm"], d["CONT"]["ln"],
                                  d["CONT"]["lt"]):
                assert w == (tuple(k) in cache)
                assert (a, b) == (cache[tuple(k)] if w else (0.0, 0.0))
        s = d["BODY"]
        for k in range(1, 56):  # the boxes (slot 0 is the floor)
            c, sn = np.cos(s["theta"][k]), np.sin(s["theta"][k])
            x = s["px"][k] + c * box[:, 0] - sn * box[:, 1]
            y = s["py"][k] + sn * box[:, 0] + c * box[:, 1]
            a = d["AABB"][k]
            assert abs(a["minx"] - x.min()) < 1e-12 and abs(a["maxx"] - x.max()) < 1e-12
            assert abs(a["miny"] - y.min()) < 1e-12 and abs(a["maxy"] - y.max()) < 1e-12
    assert pos == len(data)
