"""Seeded scene generators for the contact-physics workloads.

A scene is a list of worlds; a world is a `WorldDesc` of structure-of-arrays
body state plus CSR hull vertices in local space, counter-clockwise and
centred on the centroid.  Body order is the world's slot order (the
reference's `fromList` order, shapes/src/Physics/World.hs:118-124): the
static floor is body 0.

Random numbers come from numpy's Philox counter-based generator, so a seed
gives the same scene on every machine.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

from typing import Optional

import numpy as np


@dataclass
class WorldDesc:
    """Structure-of-arrays description of one world.

    mass/inertia of 0 mean infinite (the reference's `toInvMass2`,
    shapes/src/Physics/Constraint.hs:79-83).  `radius` (None: every body is
    a hull) makes body i a circle of radius[i] > 0 (makeCircle,
    Engine.hs:53-54), whose hull range is then empty.
    """

    px: np.ndarray
    py: np.ndarray
    theta: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    omega: np.ndarray
    mass: np.ndarray
    inertia: np.ndarray
    mu: np.ndarray
    bounce: np.ndarray
    hull_off: np.ndarray  # int32, n + 1
    hull_x: np.ndarray
    hull_y: np.ndarray
    radius: Optional[np.ndarray] = None
    capacity: int = 0  # slots (World.new capacity); 0 = n

    @property
    def n(self) -> int:
        return int(self.px.shape[0])

    @property
    def n_dynamic(self) -> int:
        return int(np.count_nonzero((self.mass != 0) | (self.inertia != 0)))

    def copy(self) -> "WorldDesc":
        return WorldDesc(**{k: (v.copy() if isinstance(v, np.ndarray) else v)
                            for k, v in self.__dict__.items()})


@dataclass
class SceneConfig:
    """Engine parameters: the reference's EngineConfig + ContactBehavior
    (shapes/src/Physics/Engine/Main.hs:34-37, Contact/Types.hs:20-25), the
    constant acceleration External, and the sweep count the reference
    hard-codes to 2 (Engine/Main.hs:81-82)."""

    dt: float = 1.0 / 60.0
    baumgarte: float = 0.01
    slop: float = 0.02
    gx: float = 0.0
    gy: float = -9.8
    iterations: int = 4
    steps: int = 300


@dataclass
class Scene:
    worlds: list
    config: SceneConfig
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def n_bodies(self) -> int:
        return sum(w.n for w in self.worlds)

    @property
    def n_dynamic(self) -> int:
        return sum(w.n_dynamic for w in self.worlds)


# ---------------------------------------------------------------- geometry

def rectangle_vertices(w: float, h: float):
    """rectangleVertices (shapes/src/Physics/Contact/ConvexHull.hs:135-145)."""
    w2 = w / 2.0
    h2 = h / 2.0
    return [(w2, h2), (-w2, h2), (-w2, -h2), (w2, -h2)]


def polygon_mass_properties(xs, ys, density=1.0):
    """Area, centroid and second moment about the centroid of a CCW polygon
    (closed-form shoelace sums)."""
    x = np.asarray(xs, dtype=np.float64)
    y = np.asarray(ys, dtype=np.float64)
    x1 = np.roll(x, -1)
    y1 = np.roll(y, -1)
    cr = x * y1 - x1 * y
    area = 0.5 * cr.sum()
    cx = ((x + x1) * cr).sum() / (6.0 * area)
    cy = ((y + y1) * cr).sum() / (6.0 * area)
    i_origin = (cr * (x * x + x * x1 + x1 * x1 + y * y + y * y1 + y1 * y1)).sum() / 12.0
    mass = density * area
    inertia = density * i_origin - mass * (cx * cx + cy * cy)
    return mass, (cx, cy), inertia


def random_polygon(rng: np.random.Generator, radius_lo=0.3, radius_hi=1.0):
    """Convex 3-8-gon: vertices on a circle of radius U[lo,hi] at angles
    2*pi*k/n + U(-0.25,0.25)*(2*pi/n), re-centred on the centroid."""
    n = int(rng.integers(3, 9))
    step = 2.0 * math.pi / n
    ang = np.arange(n) * step + rng.uniform(-0.25, 0.25, n) * step
    ang.sort()
    r = rng.uniform(radius_lo, radius_hi)
    xs = r * np.cos(ang)
    ys = r * np.sin(ang)
    _, (cx, cy), _ = polygon_mass_properties(xs, ys)
    return list(zip((xs - cx).tolist(), (ys - cy).tolist()))


class _Builder:
    def __init__(self):
        self.cols = {k: [] for k in ("px", "py", "theta", "vx", "vy", "omega",
                                     "mass", "inertia", "mu", "bounce")}
        self.off = [0]
        self.hx = []
        self.hy = []
        self.radius = []

    def add(self, verts, pos, vel=(0.0, 0.0), omega=0.0, theta=0.0,
            mass=0.0, inertia=0.0, mu=0.5, bounce=0.0):
        c = self.cols
        c["px"].append(pos[0]); c["py"].append(pos[1]); c["theta"].append(theta)
        c["vx"].append(vel[0]); c["vy"].append(vel[1]); c["omega"].append(omega)
        c["mass"].append(mass); c["inertia"].append(inertia)
        c["mu"].append(mu); c["bounce"].append(bounce)
        for (x, y) in verts:
            self.hx.append(x)
            self.hy.append(y)
        self.off.append(len(self.hx))
        self.radius.append(0.0)

    def add_circle(self, radius, pos, **kw):
        """A circle of `radius` (makeCircle, Engine.hs:53-54)."""
        self.add([], pos, **kw)
        self.radius[-1] = float(radius)

    def add_dynamic(self, verts, pos, density=1.0, **kw):
        m, _, inertia = polygon_mass_properties([v[0] for v in verts],
                                                [v[1] for v in verts], density)
        self.add(verts, pos, mass=m, inertia=inertia, **kw)

    def build(self) -> WorldDesc:
        arr = {k: np.asarray(v, dtype=np.float64) for k, v in self.cols.items()}
        rad = np.asarray(self.radius, dtype=np.float64)
        return WorldDesc(hull_off=np.asarray(self.off, dtype=np.int32),
                         hull_x=np.asarray(self.hx, dtype=np.float64),
                         hull_y=np.asarray(self.hy, dtype=np.float64),
                         radius=rad if (rad > 0).any() else None, **arr)


# ------------------------------------------------------------------ scenes

def pyramid(seed: int = 1001, rows: int = 10, steps: int = 300,
            iterations: int = 4) -> Scene:
    """Pyramid of `rows` rows of 1x1 boxes (55 for 10 rows), mass 1,
    x-jitter U(-0.01, 0.01), on a static 40x1 floor; g = -9.8, dt = 1/60,
    friction 0.5, Baumgarte/slop 0.01/0.02 (shapes/src/Physics/Scenes/
    Stacks.hs:87-88)."""
    rng = np.random.Generator(np.random.Philox(seed))
    b = _Builder()
    b.add(rectangle_vertices(40.0, 1.0), (0.0, -0.5), mu=0.5)
    box = rectangle_vertices(1.0, 1.0)
    for r in range(rows):
        for c in range(rows - r):
            x = (c - (rows - 1 - r) / 2.0) * 1.0 + rng.uniform(-0.01, 0.01)
            b.add(box, (x, 0.5 + r), mass=1.0, inertia=1.0 / 6.0, mu=0.5)
    cfg = SceneConfig(dt=1.0 / 60.0, gy=-9.8, iterations=iterations, steps=steps)
    return Scene([b.build()], cfg, name=f"pyramid{rows}",
                 meta={"seed": seed})


def reference_stacks(dims=(30, 30), spacing=0.0, steps: int = 10) -> Scene:
    """The reference benchmark scene Stacks.makeScene (30,30) 0 with
    runWorld 0.01 (shapes/bench/Main.hs:19-26; shapes/src/Physics/Scenes/
    Stacks.hs:13-113): 0.2x0.2 boxes of (mass, inertia) = (2, 1) in columns
    left to right, bottom to top, on a static 18x1 floor at (0,-6);
    constant acceleration (0,-2); ContactBehavior 0.01 0.02; 2 sweeps.
    Positions accumulate with the same floating-point additions as the
    Haskell `iterate (+ w)` and `_2 %~ (+ (h + spacing))`."""
    n_w, n_h = dims
    w = h = 0.2
    b = _Builder()
    b.add(rectangle_vertices(18.0, 1.0), (0.0, -6.0), mu=0.2)
    box = rectangle_vertices(w, h)
    center, bottom = 0.0, -4.5
    leftmost = center - (w * float(n_w - 1) / 2.0)
    left = leftmost
    for _ in range(n_w):
        y = bottom
        for _ in range(n_h):
            b.add(box, (left, y), mass=2.0, inertia=1.0, mu=0.2)
            y = y + (h + spacing)
        left = left + w
    cfg = SceneConfig(dt=0.01, gy=-2.0, iterations=2, steps=steps)
    return Scene([b.build()], cfg, name="reference_stacks_30x30",
                 meta={"dims": dims, "spacing": spacing})


def balls(dims=(10, 10), diameter: float = 0.5, spacing: float = 0.0, steps: int = 300,
          iterations: int = 2) -> Scene:
    """The reference's circle scene Balls.makeScene dims diameter spacing
    (shapes/src/Physics/Scenes/Balls.hs:16-76): on the static 18x1 floor at
    (0, -6) (boxFloor', Stacks.hs:19-32), `dims` = (columns, rows) stacks
    side by side from x = -(diameter * (columns - 1) / 2), alternating a
    column of circles of diameter `diameter` (mass (2, 1), mu 0.2, bounce
    0.5: circle', Balls.hs:16-22) and a column of diameter x diameter boxes
    (mass (2, 1), mu 0.2, bounce 0: box', Stacks.hs:22-28), bottoms at
    y = -4.5 stepping up by diameter + spacing; constant acceleration
    (0, -2), ContactBehavior 0.01 0.02 (Stacks.hs:84-88).  Positions
    accumulate with the same additions as the Haskell `iterate`."""
    n_w, n_h = dims
    b = _Builder()
    b.add(rectangle_vertices(18.0, 1.0), (0.0, -6.0), mu=0.2)
    left = 0.0 - (diameter * float(n_w - 1) / 2.0)
    circle = True
    for _ in range(n_w):
        y = -4.5
        for _ in range(n_h):
            if circle:
                b.add_circle(diameter / 2.0, (left, y), mass=2.0, inertia=1.0, mu=0.2, bounce=0.5)
            else:
                b.add(rectangle_vertices(diameter, diameter), (left, y), mass=2.0, inertia=1.0,
                      mu=0.2)
            y = y + (diameter + spacing)
        left = left + diameter
        circle = not circle
    cfg = SceneConfig(dt=0.01, gy=-2.0, iterations=iterations, steps=steps)
    return Scene([b.build()], cfg, name=f"balls_{n_w}x{n_h}",
                 meta={"dims": dims, "diameter": diameter, "spacing": spacing})


def circle_world(seed: int, count: int = 60, cols: int = 10, spacing: float = 1.2,
                 circle_frac: float = 0.5) -> WorldDesc:
    """Circles (radius U[0.2, 0.6], density 1) mixed with random 3-8-gons on
    a grid above a static floor, with seeded velocities."""
    rng = np.random.Generator(np.random.Philox(seed))
    b = _Builder()
    b.add(rectangle_vertices(cols * spacing + 10.0, 1.0), (0.0, -0.5), mu=0.5)
    x0 = -(cols - 1) * spacing / 2.0
    for k in range(count):
        r, c = divmod(k, cols)
        pos = (x0 + c * spacing + rng.uniform(-0.1, 0.1), 1.0 + r * spacing)
        vel = (rng.uniform(-1.0, 1.0), rng.uniform(-1.0, 1.0))
        if rng.uniform() < circle_frac:
            rad = rng.uniform(0.2, 0.6)
            m = np.pi * rad * rad
            b.add_circle(rad, pos, vel=vel, mass=m, inertia=0.5 * m * rad * rad,
                         mu=rng.uniform(0.2, 0.8), bounce=rng.uniform(0.0, 0.5))
        else:
            b.add_dynamic(random_polygon(rng), pos, vel=vel, mu=rng.uniform(0.2, 0.8))
    return b.build()


def _polygon_field(rng, b: _Builder, count: int, cols: int, spacing: float,
                   y0: float, square_frac: float = 0.0, mu_range=(0.2, 0.8)):
    x0 = -(cols - 1) * spacing / 2.0
    for k in range(count):
        r, c = divmod(k, cols)
        if square_frac > 0.0 and rng.uniform() < square_frac:
            side = rng.uniform(0.5, 1.0)
            verts = rectangle_vertices(side, side)
        else:
            verts = random_polygon(rng)
        pos = (x0 + c * spacing, y0 + r * spacing)
        vel = (rng.uniform(-1.0, 1.0), rng.uniform(-1.0, 1.0))
        b.add_dynamic(verts, pos, vel=vel, mu=rng.uniform(*mu_range))


def polygon_pile(seed: int = 1002, count: int = 8192, steps: int = 600,
                 iterations: int = 8, cols: int = 90,
                 floor_width: float = 200.0) -> Scene:
    """One large world: `count` random convex 3-8-gons (density 1) on a
    seeded grid at spacing 2.1 (`cols` columns, starting at y = 2) with
    v ~ U[-1,1]^2 and mu ~ U[0.2,0.8], dropped onto a static floor."""
    rng = np.random.Generator(np.random.Philox(seed))
    b = _Builder()
    b.add(rectangle_vertices(floor_width, 1.0), (0.0, -0.5), mu=0.5)
    _polygon_field(rng, b, count, cols, 2.1, 2.0)
    cfg = SceneConfig(dt=1.0 / 60.0, gy=-9.8, iterations=iterations, steps=steps)
    return Scene([b.build()], cfg, name=f"polygon_pile{count}",
                 meta={"seed": seed})


def world_batch(seed: int = 1003, n_worlds: int = 4096, lo: int = 8,
                hi: int = 1024, steps: int = 300, iterations: int = 6,
                dt: float = 1.0 / 60.0) -> Scene:
    """Independent worlds with log-uniform body counts in [lo, hi]; half
    squares (side U[0.5,1]) and half random 3-8-gons, each world on its own
    static floor, seeded initial velocities."""
    rng = np.random.Generator(np.random.Philox(seed))
    worlds = []
    counts = np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), n_worlds)))
    for n in counts.astype(int):
        cols = max(2, int(math.ceil(math.sqrt(n))))
        b = _Builder()
        b.add(rectangle_vertices(cols * 2.1 + 10.0, 1.0), (0.0, -0.5), mu=0.5)
        _polygon_field(rng, b, int(n), cols, 2.1, 2.0, square_frac=0.5)
        worlds.append(b.build())
    cfg = SceneConfig(dt=dt, gy=-9.8, iterations=iterations, steps=steps)
    return Scene(worlds, cfg, name=f"world_batch{n_worlds}",
                 meta={"seed": seed})


def random_world(seed: int, count: int = 40, cols: int = 8,
                 spacing: float = 1.6, square_frac: float = 0.5) -> WorldDesc:
    """Small mixed world for parity tests (tighter spacing -> early contacts)."""
    rng = np.random.Generator(np.random.Philox(seed))
    b = _Builder()
    b.add(rectangle_vertices(cols * spacing + 6.0, 1.0), (0.0, -0.5), mu=0.5)
    _polygon_field(rng, b, count, cols, spacing, 1.2, square_frac=square_frac)
    return b.build()


# ----------------------------------------------------- vectorized generators

def polygon_arrays(rng: np.random.Generator, m: int, square_frac: float = 0.0,
                   radius_lo: float = 0.3, radius_hi: float = 1.0):
    """m convex bodies at once: random 3-8-gons as in random_polygon (and a
    fraction of squares with side U[0.5, 1]).  Returns (counts, xs, ys,
    mass, inertia) with xs/ys padded to 8 columns, vertices CCW and centred
    on the centroid, density 1."""
    n = rng.integers(3, 9, m)
    is_sq = rng.uniform(size=m) < square_frac
    n = np.where(is_sq, 4, n)
    k = np.arange(8)[None, :]
    step = 2.0 * np.pi / n[:, None]
    ang = k * step + rng.uniform(-0.25, 0.25, (m, 8)) * step
    ang = np.where(k < n[:, None], ang, np.inf)
    ang.sort(axis=1)
    r = rng.uniform(radius_lo, radius_hi, m)[:, None]
    valid = k < n[:, None]
    safe = np.where(valid, ang, 0.0)
    xs = np.where(valid, r * np.cos(safe), 0.0)
    ys = np.where(valid, r * np.sin(safe), 0.0)
    side = rng.uniform(0.5, 1.0, m)[:, None] / 2.0
    sqx = np.array([1.0, -1.0, -1.0, 1.0, 0, 0, 0, 0])[None, :] * side
    sqy = np.array([1.0, 1.0, -1.0, -1.0, 0, 0, 0, 0])[None, :] * side
    xs = np.where(is_sq[:, None], sqx, xs)
    ys = np.where(is_sq[:, None], sqy, ys)
    nxt = np.where(k + 1 < n[:, None], k + 1, 0)
    x1 = np.take_along_axis(xs, nxt, axis=1)
    y1 = np.take_along_axis(ys, nxt, axis=1)
    cr = np.where(valid, xs * y1 - x1 * ys, 0.0)
    area = 0.5 * cr.sum(1)
    cx = ((xs + x1) * cr).sum(1) / (6.0 * area)
    cy = ((ys + y1) * cr).sum(1) / (6.0 * area)
    i0 = (cr * (xs * xs + xs * x1 + x1 * x1 + ys * ys + ys * y1 + y1 * y1)).sum(1) / 12.0
    inertia = i0 - area * (cx * cx + cy * cy)
    xs = np.where(valid, xs - cx[:, None], 0.0)
    ys = np.where(valid, ys - cy[:, None], 0.0)
    return n, xs, ys, area, inertia


def _world_from_arrays(n, xs, ys, mass, inertia, px, py, vx, vy, mu, floor_w,
                       floor_mu=0.5):
    m = len(n)
    fl = rectangle_vertices(floor_w, 1.0)
    counts = np.concatenate([[4], n]).astype(np.int32)
    off = np.zeros(m + 2, dtype=np.int32)
    np.cumsum(counts, out=off[1:])
    valid = np.arange(8)[None, :] < n[:, None]
    hx = np.concatenate([[v[0] for v in fl], xs[valid]])
    hy = np.concatenate([[v[1] for v in fl], ys[valid]])

    def col(first, rest):
        return np.concatenate([[first], rest]).astype(np.float64)

    z = np.zeros(m)
    return WorldDesc(px=col(0.0, px), py=col(-0.5, py), theta=col(0.0, z),
                     vx=col(0.0, vx), vy=col(0.0, vy), omega=col(0.0, z),
                     mass=col(0.0, mass), inertia=col(0.0, inertia),
                     mu=col(floor_mu, mu), bounce=col(0.0, z),
                     hull_off=off, hull_x=hx, hull_y=hy)


def grid_world(rng, count: int, cols: int, spacing: float, y0: float,
               square_frac: float, vel: float = 1.0) -> WorldDesc:
    """`count` bodies on a `cols`-wide grid (row-major from y0) above a
    static floor as wide as the grid plus 10."""
    n, xs, ys, mass, inertia = polygon_arrays(rng, count, square_frac)
    k = np.arange(count)
    r, c = np.divmod(k, cols)
    x0 = -(cols - 1) * spacing / 2.0
    px = x0 + c * spacing
    py = y0 + r * spacing
    vx = rng.uniform(-vel, vel, count)
    vy = rng.uniform(-vel, vel, count)
    mu = rng.uniform(0.2, 0.8, count)
    return _world_from_arrays(n, xs, ys, mass, inertia, px, py, vx, vy, mu,
                              floor_w=cols * spacing + 10.0)


def world_batch_fast(seed: int = 1003, n_worlds: int = 4096, lo: int = 8,
                     hi: int = 1024, steps: int = 300, iterations: int = 6,
                     dt: float = 1.0 / 60.0) -> Scene:
    """4096 independent worlds with log-uniform body counts in [lo, hi],
    half squares (side U[0.5,1]) and half random 3-8-gons, each on its own
    static floor with seeded initial velocities (vectorized generator)."""
    rng = np.random.Generator(np.random.Philox(seed))
    counts = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), n_worlds))).astype(int)
    worlds = [grid_world(rng, int(c), max(2, int(np.ceil(np.sqrt(c)))), 2.1, 2.0, 0.5)
              for c in counts]
    cfg = SceneConfig(dt=dt, gy=-9.8, iterations=iterations, steps=steps)
    return Scene(worlds, cfg, name=f"world_batch{n_worlds}", meta={"seed": seed})


def pile_batch(seed: int = 1005, n_worlds: int = 1024, lo: int = 2048,
               hi: int = 32768, layers: int = 60, steps: int = 1000,
               iterations: int = 4, dt: float = 1.0 / 120.0) -> Scene:
    """Worlds of log-uniform 2k-32k random 3-8-gons stacked in columns up to
    `layers` tall (spacing 2.05, velocities U[-0.2, 0.2]), each on its own
    static floor."""
    rng = np.random.Generator(np.random.Philox(seed))
    counts = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), n_worlds))).astype(int)
    worlds = []
    for c in counts:
        cols = int(np.ceil(c / layers))
        worlds.append(grid_world(rng, int(c), cols, 2.05, 1.05, 0.0, vel=0.2))
    cfg = SceneConfig(dt=dt, gy=-9.8, iterations=iterations, steps=steps)
    return Scene(worlds, cfg, name=f"pile_batch{n_worlds}", meta={"seed": seed})
