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
    big = engine.ConfigC(1 / 60, 0.01, 0.02, 0, -9.8, 4, engine.SOLVER_COLORED_F32, 0, ppb, 0)
    assert L.shp_batch_create(C.byref(d), 1, C.byref(big), C.byref(h)) == 4
    assert b"coloured" in L.shp_last_error()
