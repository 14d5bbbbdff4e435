"""The GPU's dumps (shp_batch_dump, drop-in FP64 order) equal the golden
fixtures byte for byte: the pyramid's at every checkpoint and after a
contact probe, the 8192-polygon pile's digests over all 600 steps; and a
probe changes nothing that the next step reads."""
import json
import os

import numpy as np
import pytest

import golden_util
from shapes_amd import engine

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_c1_dumps_equal_golden():
    with open(os.path.join(HERE, "c1_pyramid.shpd"), "rb") as f:
        want = f.read()
    sc = golden_util.c1_scene()
    gpu = engine.make_batch(sc.worlds, sc.config)
    out, done = [], 0
    for s in golden_util.C1_CHECKPOINTS:
        gpu.step(s - done)
        done = s
        out.append(gpu.dump(0))
    gpu.probe_contacts()
    out.append(gpu.dump(0))
    got = b"".join(out)
    assert len(got) == len(want)
    assert got == want
    gpu.close()


def test_c2_digests_equal_golden():
    with open(os.path.join(HERE, "c2_polygon_pile.json")) as f:
        want = json.load(f)
    sc = golden_util.c2_scene()
    print(f"golden pile: {sc.n_bodies} bodies, 600 steps", flush=True)
    gpu = engine.make_batch(sc.worlds, sc.config)
    done = 0
    for s in golden_util.C2_CHECKPOINTS:
        gpu.step(s - done)
        done = s
        st = gpu.stats()
        got = [golden_util.digest(gpu.dump(0)), int(gpu.filled(0)), int(st["pairs"]),
               int(st["contacts"])]
        assert got == want[str(s)], s
    gpu.close()


@pytest.mark.parametrize("solver", [engine.SOLVER_ORDERED_F64, engine.SOLVER_COLORED_F32])
def test_probe_has_no_side_effects(solver):
    """A batch probed before every step steps exactly like one that is not."""
    sc = golden_util.c1_scene()
    a = engine.make_batch(sc.worlds, sc.config, solver=solver)
    b = engine.make_batch(sc.worlds, sc.config, solver=solver)
    for step in range(60):
        a.probe_contacts()
        a.step(1)
        b.step(1)
        assert a.dump(0) == b.dump(0), step
    a.close()
    b.close()
