"""ctypes binding of the CPU FP64 oracle (liboracle_shapes.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker.  The product path
(shapes_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_shapes.so")

TRIG_ENGINE = 0
TRIG_LIBM = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


SOLVER_ORDERED = 0   # the reference's sequential FP64 sweep
SOLVER_COLORED = 1   # restatement of the GPU's coloured FP32 schedule
BROADPHASE_GRID = 0   # the reference's grid (Grid.hs), the default
BROADPHASE_BRUTE = 1  # every pair (Aabb.hs culledKeys)


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("baumgarte", C.c_double),
                ("slop", C.c_double), ("gx", C.c_double), ("gy", C.c_double),
                ("iterations", C.c_int32), ("trig", C.c_int32),
                ("solver", C.c_int32), ("world", C.c_int32),
                ("broadphase", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("pairs", C.c_int32), ("contacts", C.c_int32),
                ("warm_hits", C.c_int32), ("dag_levels", C.c_int32),
                ("max_depth", C.c_double), ("normal_residual", C.c_double),
                ("friction_residual", C.c_double)]


def build(force: bool = False) -> None:
    """Compile the oracle with its Makefile (gcc only)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-s"] + (["-B"] if force else []),
                       check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_engine_create.restype = C.c_void_p
        L.orc_engine_append.restype = C.c_int32
        L.orc_engine_append.argtypes = [C.c_void_p, _dp, C.c_double, C.c_int32, _dp, _dp]
        L.orc_engine_delete.restype = C.c_int32
        L.orc_engine_delete.argtypes = [C.c_void_p, C.c_int32]
        L.orc_engine_filled.restype = C.c_int32
        L.orc_engine_filled.argtypes = [C.c_void_p]
        L.orc_engine_read_slots.argtypes = [C.c_void_p, C.POINTER(C.c_uint8)]
        L.orc_engine_probe.argtypes = [C.c_void_p]
        L.orc_engine_num_cache.argtypes = [C.c_void_p]
        L.orc_engine_num_cache.restype = C.c_int32
        L.orc_engine_read_cache.argtypes = [C.c_void_p, _ip, _dp]
        L.orc_engine_create.argtypes = [C.c_int32] + [_dp] * 10 + [_ip, _dp, _dp, _dp,
                                                                  C.POINTER(Params), C.c_int32]
        L.orc_engine_destroy.argtypes = [C.c_void_p]
        L.orc_engine_step.argtypes = [C.c_void_p, C.c_int32]
        for f in ("orc_engine_read_bodies", "orc_engine_read_hulls",
                  "orc_engine_read_aabbs"):
            getattr(L, f).argtypes = [C.c_void_p, _dp]
        for f in ("orc_engine_num_vertices", "orc_engine_num_pairs",
                  "orc_engine_num_contacts"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_int32
        L.orc_engine_read_pairs.argtypes = [C.c_void_p, _ip]
        L.orc_engine_read_contacts.argtypes = [C.c_void_p, _ip, _ip, _dp, _dp, _ip]
        L.orc_engine_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
        L.orc_engine_set_cache.argtypes = [C.c_void_p, C.c_int32, _ip, _dp]
        L.orc_hull_contacts.argtypes = [C.c_int32, _dp, _dp, C.c_int32, _dp, _dp,
                                        _ip, _ip, _dp, _ip]
        L.orc_hull_contacts.restype = C.c_int32
        L.orc_shape_contacts.argtypes = [C.c_int32, _dp, _dp, _dp, C.c_int32, _dp, _dp, _dp,
                                         _ip, _ip, _dp]
        L.orc_shape_contacts.restype = C.c_int32
        L.orc_closest_simplex.argtypes = [C.c_int32, _dp, _dp, C.c_double, C.c_double, _ip]
        L.orc_closest_simplex.restype = C.c_int32
        L.orc_solve_constraint.argtypes = [_dp, C.c_double, _dp, _dp]
        L.orc_unordered_pairs.argtypes = [C.c_int32, _ip]
        L.orc_unordered_pairs.restype = C.c_int32
        L.orc_sincos.argtypes = [C.c_double, C.c_int32, _dp, _dp]
        L.orc_engines_run.argtypes = [C.POINTER(C.c_void_p), C.c_int32,
                                      C.c_int32, C.c_int32]
        L.orc_engines_run.restype = C.c_double
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


class OracleEngine:
    """One world stepped by the oracle (updateWorld semantics)."""

    def __init__(self, world, config, trig: int = TRIG_ENGINE, solver: int = SOLVER_ORDERED,
                 world_index: int = 0, broadphase: int = BROADPHASE_GRID):
        L = lib()
        keep = []
        args = []
        for name in ("px", "py", "theta", "vx", "vy", "omega", "mass",
                     "inertia", "mu", "bounce"):
            a, p = _d(getattr(world, name))
            keep.append(a)
            args.append(p)
        ho, hop = _i(world.hull_off)
        hx, hxp = _d(world.hull_x)
        hy, hyp = _d(world.hull_y)
        keep += [ho, hx, hy]
        prm = Params(config.dt, config.baumgarte, config.slop, config.gx,
                     config.gy, config.iterations, trig, solver, world_index, broadphase)
        capacity = int(getattr(world, "capacity", 0) or 0)
        self.n = max(world.n, capacity)  # slots
        rad = getattr(world, "radius", None)
        if rad is not None:
            rad, radp = _d(rad)
            keep.append(rad)
        else:
            radp = None
        self._h = L.orc_engine_create(world.n, *args, hop, hxp, hyp, radp, C.byref(prm),
                                      capacity)
        self._L = L

    def append(self, px, py, theta=0.0, vx=0.0, vy=0.0, omega=0.0, mass=1.0, inertia=1.0,
               mu=0.5, bounce=0.0, hull=None, radius=0.0) -> int:
        """World.append; returns the slot, or -1 when the world is full."""
        b10, b10p = _d([px, py, theta, vx, vy, omega, mass, inertia, mu, bounce])
        hx, hxp = _d([v[0] for v in hull] if hull is not None else [0.0])
        hy, hyp = _d([v[1] for v in hull] if hull is not None else [0.0])
        return self._L.orc_engine_append(self._h, b10p, radius,
                                         len(hull) if hull is not None else 0, hxp, hyp)

    def delete(self, slot: int) -> int:
        """World.delete (safeDelete); -1 when the slot is already empty."""
        return self._L.orc_engine_delete(self._h, slot)

    def filled(self) -> int:
        return self._L.orc_engine_filled(self._h)

    def slots(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.uint8)
        self._L.orc_engine_read_slots(self._h, out.ctypes.data_as(C.POINTER(C.c_uint8)))
        return out.astype(bool)

    def close(self):
        if self._h:
            self._L.orc_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, steps: int = 1):
        self._L.orc_engine_step(self._h, steps)

    def bodies(self) -> np.ndarray:
        out = np.empty((self.n, 6), dtype=np.float64)
        self._L.orc_engine_read_bodies(self._h, out.ctypes.data_as(_dp))
        return out

    def hulls(self) -> np.ndarray:
        nv = self._L.orc_engine_num_vertices(self._h)
        out = np.empty((nv, 4), dtype=np.float64)
        self._L.orc_engine_read_hulls(self._h, out.ctypes.data_as(_dp))
        return out

    def aabbs(self) -> np.ndarray:
        out = np.empty((self.n, 4), dtype=np.float64)
        self._L.orc_engine_read_aabbs(self._h, out.ctypes.data_as(_dp))
        return out

    def pairs(self) -> np.ndarray:
        n = self._L.orc_engine_num_pairs(self._h)
        out = np.empty((n, 2), dtype=np.int32)
        if n:
            self._L.orc_engine_read_pairs(self._h, out.ctypes.data_as(_ip))
        return out

    def contacts(self) -> dict:
        n = self._L.orc_engine_num_contacts(self._h)
        keys = np.empty((n, 4), dtype=np.int32)
        flip = np.empty(n, dtype=np.int32)
        geom = np.empty((n, 5), dtype=np.float64)
        lam = np.empty((n, 2), dtype=np.float64)
        warm = np.empty(n, dtype=np.int32)
        if n:
            self._L.orc_engine_read_contacts(
                self._h, keys.ctypes.data_as(_ip), flip.ctypes.data_as(_ip),
                geom.ctypes.data_as(_dp), lam.ctypes.data_as(_dp),
                warm.ctypes.data_as(_ip))
        return {"keys": keys, "flip": flip, "geom": geom, "lam": lam,
                "warm": warm}

    def probe(self):
        """worldContacts: pairs and contacts of the world as it stands."""
        self._L.orc_engine_probe(self._h)

    def cache(self):
        """(keys n x 4, lam n x 2) of the warm-start cache."""
        n = self._L.orc_engine_num_cache(self._h)
        keys = np.empty((n, 4), dtype=np.int32)
        lam = np.empty((n, 2), dtype=np.float64)
        if n:
            self._L.orc_engine_read_cache(self._h, keys.ctypes.data_as(_ip),
                                          lam.ctypes.data_as(_dp))
        return keys, lam

    def stats(self) -> dict:
        s = Stats()
        self._L.orc_engine_stats(self._h, C.byref(s))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def set_cache(self, keys, lam):
        k, kp = _i(keys)
        l, lp = _d(lam)
        self._L.orc_engine_set_cache(self._h, len(k), kp, lp)


def hull_contacts(a_verts, b_verts):
    """SAT + clipping between two world-space hulls; returns (sep, list of
    (fa, fb, flip, nx, ny, px, py, depth))."""
    L = lib()
    ax, axp = _d([v[0] for v in a_verts])
    ay, ayp = _d([v[1] for v in a_verts])
    bx, bxp = _d([v[0] for v in b_verts])
    by, byp = _d([v[1] for v in b_verts])
    keys = np.zeros(4, np.int32)
    flip = np.zeros(2, np.int32)
    geom = np.zeros(10, np.float64)
    sep = np.zeros(1, np.int32)
    n = L.orc_hull_contacts(len(ax), axp, ayp, len(bx), bxp, byp,
                            keys.ctypes.data_as(_ip), flip.ctypes.data_as(_ip),
                            geom.ctypes.data_as(_dp), sep.ctypes.data_as(_ip))
    out = []
    for k in range(n):
        out.append((int(keys[2 * k]), int(keys[2 * k + 1]), int(flip[k]),
                    *[float(g) for g in geom[5 * k:5 * k + 5]]))
    return bool(sep[0]), out


def _shape_args(shape):
    """(n, xs, ys, circle) of a shape: a list of world vertices, or a
    ("circle", radius, (cx, cy)) tuple."""
    if isinstance(shape, tuple) and shape and shape[0] == "circle":
        return 0, [0.0], [0.0], [float(shape[1]), float(shape[2][0]), float(shape[2][1])]
    return len(shape), [v[0] for v in shape], [v[1] for v in shape], [0.0, 0.0, 0.0]


def shape_contacts(a, b):
    """Contact.generateContacts for two shapes (hull vertex lists or
    ("circle", r, (x, y))): list of (fa, fb, flip, nx, ny, px, py, depth)."""
    L = lib()
    na, ax, ay, ac = _shape_args(a)
    nb, bx, by, bc = _shape_args(b)
    arrs = [_d(v) for v in (ax, ay, ac, bx, by, bc)]
    keys = np.zeros(4, np.int32)
    flip = np.zeros(2, np.int32)
    geom = np.zeros(10, np.float64)
    n = L.orc_shape_contacts(na, arrs[0][1], arrs[1][1], arrs[2][1], nb, arrs[3][1],
                             arrs[4][1], arrs[5][1], keys.ctypes.data_as(_ip),
                             flip.ctypes.data_as(_ip), geom.ctypes.data_as(_dp))
    return [(int(keys[2 * k]), int(keys[2 * k + 1]), int(flip[k]),
             *[float(g) for g in geom[5 * k:5 * k + 5]]) for k in range(n)]


def closest_simplex(verts, origin):
    """GJK.closestSimplex of a world-space hull to a point: vertex indices,
    most recently added first."""
    L = lib()
    x, xp = _d([v[0] for v in verts])
    y, yp = _d([v[1] for v in verts])
    idx = np.zeros(3, np.int32)
    n = L.orc_closest_simplex(len(verts), xp, yp, float(origin[0]), float(origin[1]),
                              idx.ctypes.data_as(_ip))
    return idx[:n].tolist()


def solve_constraint(j, b, vel6, im4):
    L = lib()
    jj, jp = _d(j)
    v = np.ascontiguousarray(vel6, dtype=np.float64).copy()
    m, mp = _d(im4)
    L.orc_solve_constraint(jp, float(b), v.ctypes.data_as(_dp), mp)
    return v


def unordered_pairs(n: int) -> np.ndarray:
    L = lib()
    out = np.empty((max(n * (n - 1) // 2, 1), 2), dtype=np.int32)
    k = L.orc_unordered_pairs(n, out.ctypes.data_as(_ip))
    return out[:k]


def sincos(x: float, trig: int = TRIG_ENGINE):
    L = lib()
    s = C.c_double()
    c = C.c_double()
    L.orc_sincos(float(x), trig, C.byref(s), C.byref(c))
    return s.value, c.value


def run_engines(engines, steps: int, threads: int = 1) -> float:
    """Step several oracle engines (independent worlds); returns seconds."""
    L = lib()
    arr = (C.c_void_p * len(engines))(*[e._h for e in engines])
    return L.orc_engines_run(arr, len(engines), steps, threads)
