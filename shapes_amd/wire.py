"""The dump format of shp_batch_dump: one world's bodies, AABBs, pairs,
contacts and warm-start cache as a little-endian byte stream (the data the
demo's overlays draw, shapes-demo/src/Physics/Demo/IOWorld.hs:74-90, and
the format of the golden fixtures).

    header  "SHPD", u16 version, u16 sections, i32 world, i32 0, i64 steps
    section char tag[4], u32 record bytes, u64 records, records...

`read` parses one or more concatenated dumps; `write` builds a dump from
arrays with the same layout, so that dumps of any engine that reads back
the same quantities compare byte for byte.
"""
from __future__ import annotations

import struct

import numpy as np

VERSION = 1
BODIES, AABBS, PAIRS, CONTACTS, CACHE = 1, 2, 4, 8, 16
ALL = 31

RECORDS = {
    "BODY": np.dtype([("slot", "<i4"), ("pad", "<i4"), ("px", "<f8"), ("py", "<f8"),
                      ("theta", "<f8"), ("vx", "<f8"), ("vy", "<f8"), ("omega", "<f8")]),
    "AABB": np.dtype([("slot", "<i4"), ("pad", "<i4"), ("minx", "<f8"), ("maxx", "<f8"),
                      ("miny", "<f8"), ("maxy", "<f8")]),
    "PAIR": np.dtype([("i", "<i4"), ("j", "<i4")]),
    "CONT": np.dtype([("i", "<i4"), ("j", "<i4"), ("fi", "u1"), ("fj", "u1"), ("flip", "u1"),
                      ("warm", "u1"), ("pad", "<u4"), ("nx", "<f8"), ("ny", "<f8"),
                      ("px", "<f8"), ("py", "<f8"), ("depth", "<f8"), ("ln", "<f8"),
                      ("lt", "<f8")]),
    "CACH": np.dtype([("i", "<i4"), ("j", "<i4"), ("fi", "<i4"), ("fj", "<i4"), ("ln", "<f8"),
                      ("lt", "<f8")]),
}
ORDER = [("BODY", BODIES), ("AABB", AABBS), ("PAIR", PAIRS), ("CONT", CONTACTS),
         ("CACH", CACHE)]
_HEADER = struct.Struct("<4sHHiiq")
_SECTION = struct.Struct("<4sIQ")


def read(data: bytes) -> list:
    """Every dump in `data`: dicts with 'world', 'steps' and one structured
    array per section present ('BODY', 'AABB', 'PAIR', 'CONT', 'CACH')."""
    out, pos = [], 0
    while pos < len(data):
        magic, version, n_sec, world, _, steps = _HEADER.unpack_from(data, pos)
        if magic != b"SHPD" or version != VERSION:
            raise ValueError(f"not a version-{VERSION} dump at byte {pos}")
        pos += _HEADER.size
        d = {"world": world, "steps": steps}
        for _ in range(n_sec):
            tag, rec, count = _SECTION.unpack_from(data, pos)
            pos += _SECTION.size
            tag = tag.decode()
            dt = RECORDS[tag]
            if rec != dt.itemsize:
                raise ValueError(f"{tag} records of {rec} bytes, expected {dt.itemsize}")
            d[tag] = np.frombuffer(data, dtype=dt, count=count, offset=pos).copy()
            pos += rec * count
        out.append(d)
    return out


def _records(tag, n, **cols):
    a = np.zeros(n, dtype=RECORDS[tag])
    for k, v in cols.items():
        a[k] = v
    return a


def write(world: int, steps: int, bodies=None, aabbs=None, pairs=None, contacts=None,
          cache=None) -> bytes:
    """A dump from arrays: bodies (slots, n x 6 state) of the filled slots,
    aabbs (slots, n x 4), pairs (n x 2), contacts (dict with 'keys' n x 4,
    'flip', 'warm', 'geom' n x 5, 'lam' n x 2), cache (keys n x 4, lam n x 2)."""
    secs = []
    if bodies is not None:
        slots, st = bodies
        st = np.asarray(st, dtype=np.float64).reshape(-1, 6)
        secs.append(("BODY", _records("BODY", len(slots), slot=slots, px=st[:, 0], py=st[:, 1],
                                      theta=st[:, 2], vx=st[:, 3], vy=st[:, 4],
                                      omega=st[:, 5])))
    if aabbs is not None:
        slots, bx = aabbs
        bx = np.asarray(bx, dtype=np.float64).reshape(-1, 4)
        secs.append(("AABB", _records("AABB", len(slots), slot=slots, minx=bx[:, 0],
                                      maxx=bx[:, 1], miny=bx[:, 2], maxy=bx[:, 3])))
    if pairs is not None:
        p = np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
        secs.append(("PAIR", _records("PAIR", len(p), i=p[:, 0], j=p[:, 1])))
    if contacts is not None:
        k = np.asarray(contacts["keys"]).reshape(-1, 4)
        g = np.asarray(contacts["geom"], dtype=np.float64).reshape(-1, 5)
        lam = np.asarray(contacts["lam"], dtype=np.float64).reshape(-1, 2)
        secs.append(("CONT", _records("CONT", len(k), i=k[:, 0], j=k[:, 1], fi=k[:, 2],
                                      fj=k[:, 3], flip=contacts["flip"], warm=contacts["warm"],
                                      nx=g[:, 0], ny=g[:, 1], px=g[:, 2], py=g[:, 3],
                                      depth=g[:, 4], ln=lam[:, 0], lt=lam[:, 1])))
    if cache is not None:
        k, lam = cache
        k = np.asarray(k).reshape(-1, 4)
        lam = np.asarray(lam, dtype=np.float64).reshape(-1, 2)
        secs.append(("CACH", _records("CACH", len(k), i=k[:, 0], j=k[:, 1], fi=k[:, 2],
                                      fj=k[:, 3], ln=lam[:, 0], lt=lam[:, 1])))
    parts = [_HEADER.pack(b"SHPD", VERSION, len(secs), world, 0, steps)]
    for tag, a in secs:
        parts.append(_SECTION.pack(tag.encode(), a.dtype.itemsize, len(a)))
        parts.append(a.tobytes())
    return b"".join(parts)
