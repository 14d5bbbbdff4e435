"""Host-side mirror of the reference engine API over the gfx950 C ABI.

The reference drives a world with Physics.Engine.Main
(shapes/src/Physics/Engine/Main.hs:32-98): `initEngine` builds the engine
state from a `Scene`, `updateWorld` advances it one step, `stepWorld n`
repeats it and `runWorld dt scene n` does all of it.  This module keeps those
names and meanings; the state lives on the GPU behind an opaque handle of
libshapes_gpu.so (include/shapes_gpu.h).  There is no CPU fallback: if the
library is missing or no GPU is present, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from .scenes import Scene, SceneConfig, WorldDesc

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libshapes_gpu.so")

SOLVER_ORDERED_F64 = 0
SOLVER_COLORED_F32 = 1
WARM_ZERO = 0
WARM_LEARNED = 1
BROADPHASE_AUTO = 0
BROADPHASE_SCAN = 1
BROADPHASE_GRID = 2
BROADPHASE_GRID_1 = 3   # the grid, one lane per body
BROADPHASE_GRID_16 = 4  # the grid, 16 lanes per body

STATUS = {0: "OK", 1: "E_INVAL", 2: "E_OOM", 3: "E_HIP", 4: "E_UNSUPPORTED",
          5: "E_OVERFLOW", 6: "E_NONFINITE"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_fp = C.POINTER(C.c_float)


class ShapesError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class WorldDescC(C.Structure):
    _fields_ = [("n_bodies", C.c_int32)] + [
        (n, _dp) for n in ("px", "py", "theta", "vx", "vy", "omega", "mass",
                           "inertia", "mu", "bounce")
    ] + [("hull_off", _ip), ("hull_x", _dp), ("hull_y", _dp), ("radius", _dp),
         ("capacity", C.c_int32)]


class BodyDescC(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("px", "py", "theta", "vx", "vy", "omega", "mass",
                                          "inertia", "mu", "bounce", "radius")] + [
        ("n_vertices", C.c_int32), ("hull_x", _dp), ("hull_y", _dp)]


class ConfigC(C.Structure):
    _fields_ = [("dt", C.c_double), ("baumgarte", C.c_double),
                ("slop", C.c_double), ("gravity_x", C.c_double),
                ("gravity_y", C.c_double), ("iterations", C.c_int32),
                ("solver", C.c_int32), ("warm_start", C.c_int32),
                ("max_pairs_per_body", C.c_int32), ("broadphase", C.c_int32),
                ("solver_threads", C.c_int32)]


class StepStatsC(C.Structure):
    _fields_ = [("pairs", C.c_int64), ("contacts", C.c_int64),
                ("warm_hits", C.c_int64), ("learned_starts", C.c_int64),
                ("misses", C.c_int64),
                ("max_levels", C.c_int32), ("flags", C.c_int32),
                ("max_depth", C.c_double), ("normal_residual", C.c_double),
                ("friction_residual", C.c_double), ("run_max_depth", C.c_double),
                ("run_normal_residual", C.c_double),
                ("run_friction_residual", C.c_double), ("mean_max_depth", C.c_double),
                ("mean_normal_residual", C.c_double), ("mean_friction_residual", C.c_double),
                ("manifolds", C.c_int64)]


class PlanC(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "bodies", "vertices", "worlds", "max_world_bodies", "pair_capacity",
        "contact_capacity", "miss_capacity", "grid_buckets", "device_bytes",
        "max_workitems", "max_lds_bytes")] + [
        (n, C.c_int32) for n in ("launches", "broadphase", "solver_threads", "reserved")]


class LaunchC(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("workgroups", C.c_int64),
                ("workgroup_size", C.c_int32), ("lds_bytes", C.c_int32)]


EXPORTS = [
    "shp_batch_create", "shp_batch_destroy", "shp_batch_step",
    "shp_batch_step_async", "shp_batch_sync", "shp_batch_num_bodies",
    "shp_batch_read_bodies", "shp_batch_read_pairs", "shp_batch_read_contacts",
    "shp_batch_read_hulls", "shp_batch_last_stats", "shp_batch_set_cache",
    "shp_mlp_param_count", "shp_batch_set_mlp", "shp_last_error",
    "shp_version", "shp_batch_set_timing", "shp_batch_read_timing",
    "shp_batch_read_solve_profile",
    "shp_batch_set_logging", "shp_batch_read_log", "shp_mlp_create",
    "shp_mlp_destroy", "shp_mlp_get_params", "shp_mlp_forward",
    "shp_mlp_gradient", "shp_mlp_set_data", "shp_mlp_train", "shp_mlp_set_loss",
    "shp_batch_plan", "shp_mlp_forward_noalloc",
    "shp_batch_append", "shp_batch_delete", "shp_batch_filled", "shp_batch_read_slots",
    "shp_batch_num_vertices", "shp_batch_probe_contacts", "shp_batch_read_aabbs",
    "shp_batch_read_cache", "shp_batch_dump",
]

PHASES = ["reset", "broadphase_count", "scan_pairs", "broadphase_fill", "sat",
          "scan_contacts", "world_ranges", "contacts", "infer", "solve_setup", "sweeps",
          "solve_writeback", "finish", "cache_store", "colouring"]
# the kernel(s) of each phase (colouring: on the second stream)
PHASE_KERNELS = {
    "broadphase_count": ["k_grid_assign", "scan", "k_grid_fill", "k_grid_pairs", "k_broadphase"],
    "scan_pairs": ["k_scan_small", "k_scan_up", "k_scan_mid", "k_scan_down"],
    "broadphase_fill": ["k_grid_pairs", "k_broadphase"],
    "sat": ["k_sat"], "scan_contacts": ["k_scan_small", "k_scan_up", "k_scan_mid", "k_scan_down"],
    "world_ranges": ["k_world_ranges"], "contacts": ["k_contacts"], "infer": ["k_mlp_infer"],
    "solve_setup": [], "sweeps": ["k_sweep_colours", "k_sweep_waves", "k_solve_flow"],
    "solve_writeback": ["k_colour_unpack"], "finish": ["k_finish"],
    "cache_store": ["k_cache_store"], "colouring": ["k_colour"],
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libshapes_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ShapesError(4, f"{path} is missing; run shapes_amd/build.py")
    L = C.CDLL(path)
    L.shp_batch_create.argtypes = [C.POINTER(WorldDescC), C.c_int32,
                                   C.POINTER(ConfigC), C.POINTER(C.c_void_p)]
    L.shp_batch_destroy.argtypes = [C.c_void_p]
    L.shp_batch_step.argtypes = [C.c_void_p, C.c_int32]
    L.shp_batch_step_async.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    L.shp_batch_sync.argtypes = [C.c_void_p]
    L.shp_batch_num_bodies.argtypes = [C.c_void_p, C.c_int32, _ip]
    L.shp_batch_read_bodies.argtypes = [C.c_void_p, C.c_int32] + [_dp] * 6
    L.shp_batch_read_pairs.argtypes = [C.c_void_p, C.c_int32, _ip, _ip]
    L.shp_batch_read_contacts.argtypes = [C.c_void_p, C.c_int32, _ip, _ip, _ip,
                                          _dp, _dp, _ip]
    L.shp_batch_read_hulls.argtypes = [C.c_void_p, C.c_int32, _dp]
    L.shp_batch_last_stats.argtypes = [C.c_void_p, C.POINTER(StepStatsC)]
    L.shp_batch_set_cache.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _ip, _dp]
    L.shp_mlp_param_count.restype = C.c_int32
    L.shp_batch_set_mlp.argtypes = [C.c_void_p, _fp]
    L.shp_last_error.restype = C.c_char_p
    L.shp_version.restype = C.c_char_p
    L.shp_batch_set_logging.argtypes = [C.c_void_p, C.c_int64]
    L.shp_batch_read_log.argtypes = [C.c_void_p, _fp, _fp, C.POINTER(C.c_int64)]
    L.shp_mlp_create.argtypes = [_fp, C.POINTER(C.c_void_p)]
    L.shp_mlp_destroy.argtypes = [C.c_void_p]
    L.shp_mlp_get_params.argtypes = [C.c_void_p, _fp]
    L.shp_mlp_forward.argtypes = [C.c_void_p, _fp, C.c_int32, _fp]
    L.shp_mlp_forward_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                          C.c_void_p]
    L.shp_mlp_gradient.argtypes = [C.c_void_p, _fp, _fp, C.c_int32, _fp, _fp]
    L.shp_mlp_set_data.argtypes = [C.c_void_p, _fp, _fp, C.c_int64]
    L.shp_mlp_train.argtypes = [C.c_void_p, _ip, C.c_int32, C.c_int32, C.c_float, _fp]
    L.shp_mlp_set_loss.argtypes = [C.c_void_p, C.c_float]
    L.shp_batch_set_timing.argtypes = [C.c_void_p, C.c_int32]
    L.shp_batch_read_timing.argtypes = [C.c_void_p, _dp, _ip]
    L.shp_batch_read_solve_profile.argtypes = [C.c_void_p, _dp]
    L.shp_batch_plan.argtypes = [C.POINTER(WorldDescC), C.c_int32, C.POINTER(ConfigC),
                                 C.POINTER(PlanC), C.POINTER(LaunchC), C.c_int32]
    L.shp_batch_append.argtypes = [C.c_void_p, C.c_int32, C.POINTER(BodyDescC), _ip]
    L.shp_batch_delete.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    L.shp_batch_filled.argtypes = [C.c_void_p, C.c_int32, _ip]
    L.shp_batch_read_slots.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_uint8)]
    L.shp_batch_num_vertices.argtypes = [C.c_void_p, C.c_int32, _ip]
    L.shp_batch_probe_contacts.argtypes = [C.c_void_p]
    L.shp_batch_read_aabbs.argtypes = [C.c_void_p, C.c_int32, _ip, _ip, _dp]
    L.shp_batch_read_cache.argtypes = [C.c_void_p, C.c_int32, _ip, _ip, _dp]
    L.shp_batch_dump.argtypes = [C.c_void_p, C.c_int32, C.c_uint32, C.c_void_p, C.c_int64,
                                 C.POINTER(C.c_int64)]
    for name in EXPORTS:
        getattr(L, name)
    _lib = L
    return L


def _check(status: int):
    if status != 0:
        msg = _lib.shp_last_error().decode() if _lib else ""
        raise ShapesError(status, msg)


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _world_descs(worlds: Sequence[WorldDesc]):
    """C descriptors of the worlds and the arrays they point into."""
    keep = []
    descs = (WorldDescC * len(worlds))()
    for k, w in enumerate(worlds):
        d = descs[k]
        d.n_bodies = w.n
        for name in ("px", "py", "theta", "vx", "vy", "omega", "mass",
                     "inertia", "mu", "bounce"):
            a = np.ascontiguousarray(getattr(w, name), dtype=np.float64)
            keep.append(a)
            setattr(d, name, _dptr(a))
        ho = np.ascontiguousarray(w.hull_off, dtype=np.int32)
        hx = np.ascontiguousarray(w.hull_x, dtype=np.float64)
        hy = np.ascontiguousarray(w.hull_y, dtype=np.float64)
        keep += [ho, hx, hy]
        d.hull_off = ho.ctypes.data_as(_ip)
        d.hull_x = _dptr(hx)
        d.hull_y = _dptr(hy)
        if getattr(w, "radius", None) is not None:
            r = np.ascontiguousarray(w.radius, dtype=np.float64)
            keep.append(r)
            d.radius = _dptr(r)
        d.capacity = int(getattr(w, "capacity", 0) or 0)
    return descs, keep


def _config(config: SceneConfig, solver: int, warm_start: int, max_pairs_per_body: int,
            broadphase: int, solver_threads: int) -> ConfigC:
    return ConfigC(config.dt, config.baumgarte, config.slop, config.gx,
                   config.gy, config.iterations, solver, warm_start,
                   max_pairs_per_body, broadphase, solver_threads)


def plan(worlds: Sequence[WorldDesc], config: SceneConfig, solver: int = SOLVER_ORDERED_F64,
         warm_start: int = WARM_ZERO, max_pairs_per_body: int = 0,
         broadphase: int = BROADPHASE_AUTO, solver_threads: int = 0):
    """shp_batch_plan: (status, message, plan dict, launch list) of what
    make_batch would allocate and launch, computed on the host (no GPU)."""
    L = load_library()
    descs, keep = _world_descs(worlds)
    cfg = _config(config, solver, warm_start, max_pairs_per_body, broadphase, solver_threads)
    out = PlanC()
    cap = 64
    launches = (LaunchC * cap)()
    st = L.shp_batch_plan(descs, len(worlds), C.byref(cfg), C.byref(out), launches, cap)
    del keep
    msg = L.shp_last_error().decode() if st else ""
    p = {f: getattr(out, f) for f, _ in PlanC._fields_}
    rows = [{"name": launches[k].name.decode(), "workgroups": launches[k].workgroups,
             "workgroup_size": launches[k].workgroup_size, "lds_bytes": launches[k].lds_bytes}
            for k in range(min(out.launches, cap))]
    return st, msg, p, rows


class Batch:
    """Independent worlds stepped together on one GPU (n = 1 is the
    reference engine).  Each world keeps its own warm-start cache."""

    def __init__(self, worlds: Sequence[WorldDesc], config: SceneConfig,
                 solver: int = SOLVER_ORDERED_F64, warm_start: int = WARM_ZERO,
                 max_pairs_per_body: int = 0, broadphase: int = BROADPHASE_AUTO,
                 solver_threads: int = 0):
        L = load_library()
        self._L = L
        descs, keep = _world_descs(worlds)
        cfg = _config(config, solver, warm_start, max_pairs_per_body, broadphase,
                      solver_threads)
        h = C.c_void_p()
        _check(L.shp_batch_create(descs, len(worlds), C.byref(cfg), C.byref(h)))
        del keep
        self._h = h
        self.n_worlds = len(worlds)
        # slots per world (World.new capacity; n when no capacity is given)
        self.sizes = [max(w.n, int(getattr(w, "capacity", 0) or 0)) for w in worlds]
        self.config = config

    def close(self):
        if getattr(self, "_h", None):
            self._L.shp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- stepping (updateWorld / stepWorld)
    def step(self, steps: int = 1):
        _check(self._L.shp_batch_step(self._h, steps))

    def step_async(self, steps: int = 1, use_graph: bool = False):
        _check(self._L.shp_batch_step_async(self._h, steps, int(use_graph)))

    def sync(self):
        _check(self._L.shp_batch_sync(self._h))

    # ---- readback
    def bodies(self, world: int = 0) -> np.ndarray:
        """(n, 6) array of px, py, theta, vx, vy, omega."""
        n = self.sizes[world]
        cols = [np.empty(n, dtype=np.float64) for _ in range(6)]
        _check(self._L.shp_batch_read_bodies(self._h, world,
                                             *[_dptr(c) for c in cols]))
        return np.stack(cols, axis=1)

    def pairs(self, world: int = 0) -> np.ndarray:
        cnt = C.c_int32()
        _check(self._L.shp_batch_read_pairs(self._h, world, C.byref(cnt), None))
        out = np.empty((cnt.value, 2), dtype=np.int32)
        if cnt.value:
            _check(self._L.shp_batch_read_pairs(self._h, world, C.byref(cnt),
                                                out.ctypes.data_as(_ip)))
        return out

    def contacts(self, world: int = 0) -> dict:
        cnt = C.c_int32()
        _check(self._L.shp_batch_read_contacts(self._h, world, C.byref(cnt),
                                               None, None, None, None, None))
        n = cnt.value
        keys = np.empty((n, 4), dtype=np.int32)
        flip = np.empty(n, dtype=np.int32)
        geom = np.empty((n, 5), dtype=np.float64)
        lam = np.empty((n, 2), dtype=np.float64)
        warm = np.empty(n, dtype=np.int32)
        if n:
            _check(self._L.shp_batch_read_contacts(
                self._h, world, C.byref(cnt), keys.ctypes.data_as(_ip),
                flip.ctypes.data_as(_ip), _dptr(geom), _dptr(lam),
                warm.ctypes.data_as(_ip)))
        return {"keys": keys, "flip": flip, "geom": geom, "lam": lam,
                "warm": warm}

    def hulls(self, world: int = 0) -> np.ndarray:
        nv = C.c_int32()
        _check(self._L.shp_batch_num_vertices(self._h, world, C.byref(nv)))
        out = np.empty((nv.value, 4), dtype=np.float64)
        _check(self._L.shp_batch_read_hulls(self._h, world, _dptr(out)))
        return out

    # ---- debug overlays and dumps (shapes-demo IOWorld.hs:74-90)
    def probe_contacts(self):
        """worldContacts: pairs and contacts of the worlds as they stand (no
        step); pairs() / contacts() return them until the next step."""
        _check(self._L.shp_batch_probe_contacts(self._h))

    def aabbs(self, world: int = 0):
        """worldAabbs: (slots, n x 4 minx/maxx/miny/maxy) of the filled slots."""
        n = C.c_int32()
        _check(self._L.shp_batch_read_aabbs(self._h, world, C.byref(n), None, None))
        slots = np.empty(n.value, dtype=np.int32)
        boxes = np.empty((n.value, 4), dtype=np.float64)
        if n.value:
            _check(self._L.shp_batch_read_aabbs(self._h, world, C.byref(n),
                                                slots.ctypes.data_as(_ip), _dptr(boxes)))
        return slots, boxes

    def cache(self, world: int = 0):
        """The warm-start cache: (keys n x 4, lam n x 2), keys descending."""
        n = C.c_int32()
        _check(self._L.shp_batch_read_cache(self._h, world, C.byref(n), None, None))
        keys = np.empty((n.value, 4), dtype=np.int32)
        lam = np.empty((n.value, 2), dtype=np.float64)
        if n.value:
            _check(self._L.shp_batch_read_cache(self._h, world, C.byref(n),
                                                keys.ctypes.data_as(_ip), _dptr(lam)))
        return keys, lam

    def dump(self, world: int = 0, sections: int = 31) -> bytes:
        """shp_batch_dump: the world's dump bytes (format: shapes_amd.wire)."""
        size = C.c_int64()
        _check(self._L.shp_batch_dump(self._h, world, sections, None, 0, C.byref(size)))
        buf = C.create_string_buffer(max(size.value, 1))
        _check(self._L.shp_batch_dump(self._h, world, sections, buf, size.value,
                                      C.byref(size)))
        return buf.raw[:size.value]

    # ---- world mutation (World.append / World.delete, World.hs:77-87)
    def append(self, world: int, px: float, py: float, theta: float = 0.0, vx: float = 0.0,
               vy: float = 0.0, omega: float = 0.0, mass: float = 1.0, inertia: float = 1.0,
               mu: float = 0.5, bounce: float = 0.0, hull=None, radius: float = 0.0) -> int:
        """World.append: the body into the world's head empty slot; returns
        the slot.  `hull` is a sequence of local-space (x, y) vertices, or
        None for a circle of `radius`."""
        hx = np.ascontiguousarray([v[0] for v in hull] if hull is not None else [0.0],
                                  dtype=np.float64)
        hy = np.ascontiguousarray([v[1] for v in hull] if hull is not None else [0.0],
                                  dtype=np.float64)
        d = BodyDescC(px, py, theta, vx, vy, omega, mass, inertia, mu, bounce, radius,
                      len(hull) if hull is not None else 0, _dptr(hx), _dptr(hy))
        slot = C.c_int32()
        _check(self._L.shp_batch_append(self._h, world, C.byref(d), C.byref(slot)))
        return slot.value

    def delete(self, world: int, slot: int):
        """World.delete (EmptiesVector.safeDelete): the slot becomes empty."""
        _check(self._L.shp_batch_delete(self._h, world, slot))

    def filled(self, world: int = 0) -> int:
        n = C.c_int32()
        _check(self._L.shp_batch_filled(self._h, world, C.byref(n)))
        return n.value

    def slots(self, world: int = 0) -> np.ndarray:
        """Per slot: True when filled (EmptiesVector.read)."""
        out = np.empty(self.sizes[world], dtype=np.uint8)
        _check(self._L.shp_batch_read_slots(self._h, world,
                                            out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out.astype(bool)

    def stats(self) -> dict:
        s = StepStatsC()
        _check(self._L.shp_batch_last_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in StepStatsC._fields_}

    def set_cache(self, world: int, keys, lam):
        k = np.ascontiguousarray(keys, dtype=np.int32)
        l = np.ascontiguousarray(lam, dtype=np.float64)
        _check(self._L.shp_batch_set_cache(self._h, world, len(k),
                                           k.ctypes.data_as(_ip), _dptr(l)))

    def set_timing(self, enable: bool = True):
        _check(self._L.shp_batch_set_timing(self._h, int(enable)))

    def timing(self) -> dict:
        """Average milliseconds per step of each phase since set_timing."""
        ms = np.zeros(len(PHASES), dtype=np.float64)
        n = C.c_int32()
        _check(self._L.shp_batch_read_timing(self._h, _dptr(ms), C.byref(n)))
        k = max(n.value, 1)
        return {"steps": n.value, **{p: float(v) / k for p, v in zip(PHASES, ms)}}

    def solve_profile(self) -> list:
        """Average microseconds per timed step of the solve's sub-phases
        (world 0)."""
        us = np.zeros(5, dtype=np.float64)
        _check(self._L.shp_batch_read_solve_profile(self._h, _dptr(us)))
        return [round(float(x), 2) for x in us]

    def set_mlp(self, params: np.ndarray):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == MLP_NPARAM
        _check(self._L.shp_batch_set_mlp(self._h, p.ctypes.data_as(_fp)))

    def set_logging(self, capacity: int):
        _check(self._L.shp_batch_set_logging(self._h, int(capacity)))

    def read_log(self):
        """(features n x 14, targets n x 2) logged since set_logging."""
        n = C.c_int64(0)
        _check(self._L.shp_batch_read_log(self._h, None, None, C.byref(n)))
        f = np.empty((n.value, MLP_IN), dtype=np.float32)
        t = np.empty((n.value, 2), dtype=np.float32)
        if n.value:
            _check(self._L.shp_batch_read_log(self._h, f.ctypes.data_as(_fp),
                                              t.ctypes.data_as(_fp), C.byref(n)))
        return f[:n.value], t[:n.value]


MLP_IN = 14
MLP_NPARAM = 35202


class Mlp:
    """The warm-start network on the GPU (shp_mlp_*): BF16 matrix-core
    forward/backward, Adam on FP32 master weights."""

    def __init__(self, params: np.ndarray):
        self._L = load_library()
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == MLP_NPARAM
        h = C.c_void_p()
        _check(self._L.shp_mlp_create(p.ctypes.data_as(_fp), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.shp_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.empty(MLP_NPARAM, dtype=np.float32)
        _check(self._L.shp_mlp_get_params(self._h, out.ctypes.data_as(_fp)))
        return out

    def forward(self, feats) -> np.ndarray:
        f = np.ascontiguousarray(feats, dtype=np.float32)
        out = np.empty((len(f), 2), dtype=np.float32)
        _check(self._L.shp_mlp_forward(self._h, f.ctypes.data_as(_fp), len(f),
                                       out.ctypes.data_as(_fp)))
        return out

    def forward_noalloc(self, d_feats: int, n: int, d_out: int, stream: int = 0):
        """shp_mlp_forward_noalloc on device addresses (n x 14 features in,
        n x 2 outputs out), enqueued on `stream` (0: the handle's)."""
        _check(self._L.shp_mlp_forward_noalloc(self._h, C.c_void_p(d_feats), int(n),
                                               C.c_void_p(d_out), C.c_void_p(stream or None)))

    def gradient(self, feats, targets):
        f = np.ascontiguousarray(feats, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        g = np.empty(MLP_NPARAM, dtype=np.float32)
        loss = C.c_float()
        _check(self._L.shp_mlp_gradient(self._h, f.ctypes.data_as(_fp),
                                        t.ctypes.data_as(_fp), len(f),
                                        g.ctypes.data_as(_fp), C.byref(loss)))
        return g, loss.value

    def set_loss(self, huber_delta: float = 0.0):
        """Squared error (0) or Huber loss with this threshold."""
        _check(self._L.shp_mlp_set_loss(self._h, float(huber_delta)))

    def set_data(self, feats, targets):
        f = np.ascontiguousarray(feats, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        _check(self._L.shp_mlp_set_data(self._h, f.ctypes.data_as(_fp),
                                        t.ctypes.data_as(_fp), len(f)))

    def train(self, indices: np.ndarray, batch: int, lr: float,
              want_loss: bool = False):
        idx = np.ascontiguousarray(indices, dtype=np.int32)
        steps = idx.size // batch
        loss = np.empty(steps, dtype=np.float32) if want_loss else None
        _check(self._L.shp_mlp_train(self._h, idx.ctypes.data_as(_ip), steps, batch,
                                     float(lr),
                                     loss.ctypes.data_as(_fp) if want_loss else None))
        return loss


def make_batch(worlds, config, **kw) -> Batch:
    return Batch(worlds, config, **kw)


# ---------------------------------------------------------------------------
# Names of the reference's Physics.Engine / Physics.Engine.Main API.


@dataclass
class ContactBehavior:
    """ContactBehavior (shapes/src/Physics/Contact/Types.hs:20-25)."""
    contact_baumgarte: float = 0.0
    contact_penetration_slop: float = 0.0


@dataclass
class EngineConfig:
    """EngineConfig (Engine/Main.hs:34-37) plus the sweep count."""
    timestep: float
    contact_beh: ContactBehavior
    iterations: int = 2


def init_engine(scene: Scene, solver: int = SOLVER_ORDERED_F64,
                warm_start: int = WARM_ZERO) -> Batch:
    """initEngine (Engine/Main.hs:43-46): world + empty cache + externals."""
    return make_batch(scene.worlds, scene.config, solver=solver,
                      warm_start=warm_start)


def update_world(state: Batch) -> None:
    """updateWorld (Engine/Main.hs:71-86)."""
    state.step(1)


def step_world(state: Batch, steps: int) -> None:
    """stepWorld (Engine/Main.hs:88-89)."""
    state.step(steps)


def run_world(dt: float, scene: Scene, steps: int) -> Batch:
    """runWorld (Engine/Main.hs:97-98); returns the final state."""
    cfg = SceneConfig(**{**scene.config.__dict__, "dt": dt})
    state = make_batch(scene.worlds, cfg)
    state.step(steps)
    return state


def read_phys_obj(state: Batch, i: int, world: int = 0) -> dict:
    """readPhysObj (World.hs:93-94)."""
    b = state.bodies(world)[i]
    return {"pos": (b[0], b[1]), "rot_pos": b[2], "vel": (b[3], b[4]),
            "rot_vel": b[5]}
