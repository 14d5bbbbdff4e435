"""Dumps of the oracle in the format of shp_batch_dump (shapes_amd.wire),
the checkpoints of the golden fixtures, and the workloads they cover: the
55-box pyramid (4 sweeps, 300 steps) and the 8192-polygon pile (8 sweeps,
600 steps), both in the reference's FP64 order."""
import hashlib

import numpy as np

from shapes_amd import scenes, wire

C1_CHECKPOINTS = [1, 2, 10, 100, 300]
C2_CHECKPOINTS = [1, 50, 100, 200, 300, 400, 500, 600]


def oracle_dump(e, steps: int, world: int = 0) -> bytes:
    """Every section of one oracle world, as shp_batch_dump writes them."""
    slots = np.nonzero(e.slots())[0].astype(np.int32)
    keys, lam = e.cache()
    return wire.write(world, steps, bodies=(slots, e.bodies()[slots]),
                      aabbs=(slots, e.aabbs()[slots]), pairs=e.pairs(),
                      contacts=e.contacts(), cache=(keys, lam))


def c1_scene():
    return scenes.pyramid()


def c2_scene():
    return scenes.polygon_pile()


def c1_oracle_dumps(orc) -> bytes:
    """The pyramid's dumps after each checkpoint step, then one more at the
    last checkpoint after a contact probe of the settled world."""
    sc = c1_scene()
    e = orc.OracleEngine(sc.worlds[0], sc.config)
    out, done = [], 0
    for s in C1_CHECKPOINTS:
        e.step(s - done)
        done = s
        out.append(oracle_dump(e, done))
    e.probe()
    out.append(oracle_dump(e, done))
    e.close()
    return b"".join(out)


def digest(data: bytes) -> str:
    return hashlib.sha256(data).hexdigest()


def c2_oracle_digests(orc, upto: int = 600) -> dict:
    """{step: (sha256 of the dump, bodies, pairs, contacts)} of the pile."""
    sc = c2_scene()
    e = orc.OracleEngine(sc.worlds[0], sc.config)
    out, done = {}, 0
    for s in C2_CHECKPOINTS:
        if s > upto:
            break
        e.step(s - done)
        done = s
        d = oracle_dump(e, done)
        st = e.stats()
        out[str(s)] = [digest(d), int(e.slots().sum()), int(st["pairs"]), int(st["contacts"])]
    e.close()
    return out
