"""Pins the CPU oracle against the reference's known answers, closed forms,
brute force and invariants (CPU only).

Known answers come from the reference's own fixtures: the contact benchmark
boxes (shapes/bench/Physics/Contact/Benchmark.hs:16-27), the constraint
benchmark (shapes/bench/Physics/Constraint/Benchmark.hs:11-43) and the
unorderedPairs property (shapes/test/Physics/Broadphase/AabbSpec.hs:8-16).
"""
import math

import numpy as np
import pytest

from shapes_amd import scenes


def box(cx, cy, w, h):
    return [(x + cx, y + cy) for (x, y) in scenes.rectangle_vertices(w, h)]


# ---------------------------------------------------------------- known answers

def test_contact_benchmark_boxes(orc):
    """contact (makeOptBox 0 0 4 4) (makeOptBox 1 3 2 2): the boxes touch
    along y = 2; equal depths choose Flip, B's edge 2 is penetrated by A's
    vertices 1 (clipped to (0,2)) and 0."""
    sep, cs = orc.hull_contacts(box(0, 0, 4, 4), box(1, 3, 2, 2))
    assert not sep
    assert cs == [(1, 2, 1, 0.0, -1.0, 0.0, 2.0, 0.0),
                  (0, 2, 1, 0.0, -1.0, 2.0, 2.0, 0.0)]


def test_solve_constraint_benchmark(orc):
    """solveConstraint testConstraint testObjPair: J = (-0,-1,0,0,1,0), b = 0,
    v_a = (1,1,0), v_b = (1,-1,0), im = (1,2) each -> lambda = 1, both
    velocities become (1,0,0)."""
    v = orc.solve_constraint([-0.0, -1.0, 0.0, 0.0, 1.0, 0.0], 0.0,
                             [1, 1, 0, 1, -1, 0], [1, 2, 1, 2])
    assert v.tolist() == [1.0, 0.0, 0.0, 1.0, 0.0, 0.0]


@pytest.mark.parametrize("n", range(0, 31))
def test_unordered_pairs(orc, n):
    p = orc.unordered_pairs(n)
    assert len(p) == n * (n - 1) // 2
    want = [(x, y) for x in range(n - 1, 0, -1) for y in range(x - 1, -1, -1)]
    assert [tuple(r) for r in p.tolist()] == want


# ------------------------------------------------------------- arithmetic order

def test_lagrangian_uses_left_fold_dot(orc):
    """dotV6 is a left fold (shapes-math Template.hs:108-110).  The inputs
    are chosen so that the left fold and a pairwise sum round differently:
    the oracle must agree with the left fold."""
    j = [1.0, 1e-16, 1e-16, 0.0, 0.0, 0.0]
    v = [1.0, 1.0, 1.0, 0.0, 0.0, 0.0]
    left = ((((1.0 + 1e-16) + 1e-16) + 0.0) + 0.0) + 0.0
    pairwise = 1.0 + (1e-16 + 1e-16)
    assert left != pairwise
    im = [1.0, 0.0, 1.0, 0.0]
    out = orc.solve_constraint(j, 0.0, list(v), im)
    mc = ((((1.0 * 1.0 * 1.0 + 1e-16 * 1.0 * 1e-16) + 1e-16 * 0.0 * 1e-16)))
    lam = -(left + 0.0) / mc
    assert out[0] == 1.0 + ((1.0 * lam) * 1.0)
    assert out[2] == 1.0 + ((1e-16 * lam) * 0.0)


# ------------------------------------------------------------------- SAT / clip

def test_sat_axis_aligned_tie(orc):
    """A 2x1 box 1/8 into a 10x1 floor.  Both hulls' minimum overlaps are
    1/8, and the tie goes to Flip (eitherBranchBoth, Utils.hs:230-235): the
    box's bottom edge 2 is penetrated by the floor's top edge, clipped to
    x in [-0.75, 1.25]; the third clip is parallel (NaN -> ClipNone).
    Derived by hand from SAT.hs:152-218."""
    floor = box(0.0, -0.5, 10.0, 1.0)
    b = box(0.25, 0.5 - 0.125, 2.0, 1.0)
    sep, cs = orc.hull_contacts(floor, b)
    assert not sep
    assert cs == [(1, 2, 1, 0.0, -1.0, -0.75, 0.0, 0.125),
                  (0, 2, 1, 0.0, -1.0, 1.25, 0.0, 0.125)]


def test_sat_vertex_into_face(orc):
    """A triangle (0,-1/4), (1,1), (-1,1) poking 1/4 into the floor: the
    floor's top edge has the smallest overlap (Same), the incident edge is
    (v2, v0) because bcn < abn is false on the tie, and the face clip keeps
    only vertex 0 (ClipLeft).  Derived by hand from SAT.hs:152-218."""
    floor = box(0.0, -0.5, 10.0, 1.0)
    tri = [(0.0, -0.25), (1.0, 1.0), (-1.0, 1.0)]
    sep, cs = orc.hull_contacts(floor, tri)
    assert not sep
    assert cs == [(0, 0, 0, 0.0, 1.0, 0.0, -0.25, 0.25)]


def _rot90(verts, k):
    out = []
    for (x, y) in verts:
        for _ in range(k % 4):
            x, y = -y, x
        out.append((x, y))
    return out


@pytest.mark.parametrize("k", [1, 2, 3])
def test_sat_quarter_turn_invariance(orc, k):
    """Rotating both hulls by k*pi/2 (exact in binary) keeps the feature
    keys, orientation and depth, and rotates normals and points."""
    a = box(0.0, 0.0, 4.0, 2.0)
    b = [(x + 0.5, y + 0.6) for (x, y) in [(0.5, 0.25), (-0.5, 0.5), (-0.25, -0.5)]]
    sep0, c0 = orc.hull_contacts(a, b)
    sep1, c1 = orc.hull_contacts(_rot90(a, k), _rot90(b, k))
    assert not sep0 and sep0 == sep1 and len(c0) == len(c1) > 0
    for p, q in zip(c0, c1):
        assert p[:3] == q[:3] and p[7] == q[7]
        assert _rot90([(p[3], p[4])], k)[0] == (q[3], q[4])
        assert _rot90([(p[5], p[6])], k)[0] == (q[5], q[6])


def _brute_min_overlap(a, b):
    """Minimum over every edge normal of both hulls of the SAT overlap
    (numpy, independent of the fold order); None if any axis separates."""
    best = math.inf
    for (se, sp) in ((a, b), (b, a)):
        se = np.asarray(se)
        sp = np.asarray(sp)
        n = len(se)
        for e in range(n):
            d = se[(e + 1) % n] - se[e]
            nrm = np.array([d[1], -d[0]]) / math.hypot(d[0], d[1])
            ps = se @ nrm
            pp = sp @ nrm
            if pp.min() > ps.max() or pp.max() < ps.min():
                return None
            best = min(best, ps.max() - pp.min())
    return best


def test_sat_matches_bruteforce_min_axis(orc):
    rng = np.random.default_rng(7)
    checked = exact = 0
    for _ in range(400):
        pa = scenes.random_polygon(rng)
        pb = scenes.random_polygon(rng)
        off = rng.uniform(-1.5, 1.5, 2)
        th = rng.uniform(0, 2 * math.pi)
        c, s = math.cos(th), math.sin(th)
        b = [(c * x - s * y + off[0], s * x + c * y + off[1]) for (x, y) in pb]
        sep, cs = orc.hull_contacts(pa, b)
        want = _brute_min_overlap(pa, b)
        assert sep == (want is None)
        if want is not None and cs:
            # clipped points lie on the incident edge, never deeper than the
            # minimum overlap; the deepest equals it when the penetrating
            # vertex survives clipping
            checked += 1
            deepest = max(q[7] for q in cs)
            assert deepest <= want + 1e-9
            exact += abs(deepest - want) < 1e-9
            for q in cs:
                assert abs(math.hypot(q[3], q[4]) - 1.0) < 1e-12
    assert checked > 50 and exact > 0.8 * checked


# ----------------------------------------------------------------- broadphase

def _grid_pairs(aabbs, static):
    """Restatement of Grid.toGrid/culledKeys (Broadphase/Grid.hs:67-141)
    with the engine's fixed 20x20 axes of 1 m at -10 (Engine/Main.hs:40-41),
    including the aliasing of out-of-range cells."""
    cells = {}
    for k, (x0, x1, y0, y1) in enumerate(aabbs):
        for x in range(math.floor(x0 + 10.0), math.floor(x1 + 10.0) + 1):
            for y in range(math.floor(y0 + 10.0), math.floor(y1 + 10.0) + 1):
                cells.setdefault(x + 20 * y, {})[k] = (x0, x1, y0, y1)
    out = set()
    for sq in cells.values():
        ks = sorted(sq, reverse=True)
        for i, a in enumerate(ks):
            for b in ks[i + 1:]:
                if static[a] and static[b]:
                    continue
                A, B = sq[a], sq[b]
                if not (B[0] > A[1] or B[1] < A[0]) and not (B[2] > A[3] or B[3] < A[2]):
                    out.add((a, b))
    return sorted(out, reverse=True)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_broadphase_equals_grid_and_bruteforce(orc, seed):
    w = scenes.random_world(seed, count=60, spacing=1.2)
    cfg = scenes.SceneConfig(iterations=2)
    e = orc.OracleEngine(w, cfg)
    for _ in range(40):
        aabbs = e.aabbs()
        e.step(1)
        pairs = [tuple(p) for p in e.pairs().tolist()]
        static = [(m == 0 and i == 0) for m, i in zip(w.mass, w.inertia)]
        assert pairs == _grid_pairs(aabbs, static)
        assert pairs == sorted(pairs, reverse=True)
        assert len(set(pairs)) == len(pairs)


# ---------------------------------------------------------------- constraints

def test_one_contact_closed_form(orc):
    """lambda = -(J.v + b) / (J M^-1 J^T) for a single row (Constraint.hs:
    164-179), applied as v + M^-1 J^T lambda."""
    rng = np.random.default_rng(3)
    for _ in range(50):
        j = rng.uniform(-2, 2, 6)
        v = rng.uniform(-2, 2, 6)
        im = rng.uniform(0.1, 2, 4)
        b = rng.uniform(-1, 1)
        m6 = np.array([im[0], im[0], im[1], im[2], im[2], im[3]])
        lam = -(j @ v + b) / (j @ (m6 * j))
        out = orc.solve_constraint(j, b, v.copy(), im)
        np.testing.assert_allclose(out, v + m6 * j * lam, rtol=1e-12, atol=1e-12)
        # after the update the row is satisfied: J.v' + b = 0
        assert abs(j @ out + b) < 1e-9


def test_solver_invariants(orc):
    """After every step: lambda_n >= 0 and |lambda_t| <= mu_pair * lambda_n
    (positive / clampAbs, SolutionProcessors.hs:33-53); the static floor
    never moves."""
    w = scenes.random_world(11, count=50, spacing=1.3)
    cfg = scenes.SceneConfig(iterations=4)
    e = orc.OracleEngine(w, cfg)
    floor0 = e.bodies()[0].copy()
    for _ in range(120):
        e.step(1)
        c = e.contacts()
        lam = c["lam"]
        k = c["keys"]
        mu = (w.mu[k[:, 0]] + w.mu[k[:, 1]]) / 2.0
        assert (lam[:, 0] >= 0).all()
        assert (np.abs(lam[:, 1]) <= mu * lam[:, 0] + 1e-15).all()
        assert (e.bodies()[0] == floor0).all()


def test_momentum_conserved_between_dynamic_bodies(orc):
    """Two dynamic boxes colliding with no gravity and no static body:
    J_lin,a = -J_lin,b, so total linear momentum is conserved."""
    b = scenes._Builder()
    b.add_dynamic(scenes.rectangle_vertices(1, 1), (0.0, 0.0), vel=(1.0, 0.1))
    b.add_dynamic(scenes.rectangle_vertices(1, 2), (1.02, 0.3), vel=(-1.0, 0.0),
                  density=2.0)
    w = b.build()
    cfg = scenes.SceneConfig(gy=0.0, iterations=4, baumgarte=0.0, slop=0.0)
    e = orc.OracleEngine(w, cfg)
    p0 = w.mass @ e.bodies()[:, 3:5]
    hit = False
    for _ in range(30):
        e.step(1)
        hit |= e.stats()["contacts"] > 0
    assert hit
    np.testing.assert_allclose(w.mass @ e.bodies()[:, 3:5], p0, atol=1e-12)


def test_resting_box_normal_impulse(orc):
    """A box resting on the floor: at steady state the two contacts carry
    the gravity impulse m*g*dt per step."""
    b = scenes._Builder()
    b.add(scenes.rectangle_vertices(10, 1), (0.0, -0.5))
    b.add(scenes.rectangle_vertices(1, 1), (0.0, 0.5), mass=2.0, inertia=2.0 / 6.0)
    w = b.build()
    cfg = scenes.SceneConfig(iterations=8)
    e = orc.OracleEngine(w, cfg)
    e.step(200)
    lam = e.contacts()["lam"]
    assert len(lam) == 2
    np.testing.assert_allclose(lam[:, 0].sum(), 2.0 * 9.8 * cfg.dt, rtol=1e-3)


def test_warm_start_join_matches_lookup(orc):
    """applyCachedSlns' merge join (Descending.hs:47-71) reuses exactly the
    cached entries whose key equals a new contact's key."""
    w = scenes.random_world(5, count=30, spacing=1.3)
    cfg = scenes.SceneConfig(iterations=2)
    e = orc.OracleEngine(w, cfg)
    e.step(60)
    keys = e.contacts()["keys"]
    rng = np.random.default_rng(0)
    keep = rng.uniform(size=len(keys)) < 0.5
    extra = np.array([[1000, 999, 0, 0], [2, 1, 7, 7]], dtype=np.int32)
    ck = np.concatenate([keys[keep], extra])
    order = np.lexsort(ck.T[::-1])[::-1]
    ck = ck[order]
    lam = np.tile([[0.0, 0.0]], (len(ck), 1))
    e.set_cache(ck, lam)
    e.step(1)
    c = e.contacts()
    cached = {tuple(k) for k in ck.tolist()}
    want = [tuple(k) in cached for k in c["keys"].tolist()]
    assert c["warm"].astype(bool).tolist() == want


def test_free_flight_closed_form(orc):
    """No contacts: v_n = v0 + n*(g*dt) and pos += dt*v (applyExternal then
    advance, World.hs:156-169), with the same float operations."""
    b = scenes._Builder()
    b.add_dynamic(scenes.rectangle_vertices(1, 1), (3.0, 50.0), vel=(0.5, 2.0),
                  omega=0.3)
    w = b.build()
    cfg = scenes.SceneConfig(iterations=2)
    e = orc.OracleEngine(w, cfg)
    px, py, th, vx, vy, om = 3.0, 50.0, 0.0, 0.5, 2.0, 0.3
    for _ in range(50):
        vx = vx + 0.0 * cfg.dt
        vy = vy + cfg.gy * cfg.dt
        px = (cfg.dt * vx) + px
        py = (cfg.dt * vy) + py
        th = (cfg.dt * om) + th
    e.step(50)
    assert e.bodies()[0].tolist() == [px, py, th, vx, vy, om]


def test_coloured_restatement_tracks_fp64_pyramid(orc):
    """The oracle's restatement of the GPU's coloured FP32 schedule (not a
    reference function) finds the same contacts on the first step, keeps the
    10-row pyramid standing within 3 cm of the FP64 sequential run after 300
    steps (pyramid seeds 1-8 give 0.9-2.2 cm), and is reproducible."""
    sc = scenes.pyramid()
    a = orc.OracleEngine(sc.worlds[0], sc.config)
    b = orc.OracleEngine(sc.worlds[0], sc.config, solver=orc.SOLVER_COLORED)
    c = orc.OracleEngine(sc.worlds[0], sc.config, solver=orc.SOLVER_COLORED)
    a.step(1)
    b.step(1)
    c.step(1)
    assert np.array_equal(a.contacts()["keys"], b.contacts()["keys"])
    for _ in range(30):
        a.step(10)
        b.step(10)
        c.step(10)
    assert np.abs(a.bodies()[:, :2] - b.bodies()[:, :2]).max() < 0.03
    assert np.array_equal(b.bodies().view(np.uint64), c.bodies().view(np.uint64))
    lam = b.contacts()["lam"]
    assert (lam[:, 0] >= 0).all()


# ------------------------------------------------- contact update known answers

@pytest.mark.parametrize("name", ["warm", "bounce", "bounce_clamped", "bounce_two_sweeps"])
def test_contact_update_known_answers(orc, name):
    """Hand-derived answers of one step of the contact-benchmark boxes
    (tests/kat_scenes.py): cached lagrangians applied before the sweeps
    (useCache, Solvers/Contact.hs:100-114); both rows computed from the same
    pre-update velocities and applied normal first (improveContactSln,
    :124-143, applySln :55-65); bounceB with e > 0 (Restitution.hs:34-47);
    the friction bound from the accumulated normal impulse
    (Constraints/Contact.hs:101-110).  Friction from the post-normal
    velocity, a bound from the cached or the new normal impulse, an
    unapplied cache, or a bounce term with swapped radii each change these
    numbers."""
    import kat_scenes as K

    case = K.CASES[name]
    w = K.world(**case["world"])
    e = orc.OracleEngine(w, K.config(case["iterations"]))
    if case["cache"] is not None:
        e.set_cache(np.array(K.KEYS, dtype=np.int32), np.array(case["cache"]))
    e.step(1)
    c = e.contacts()
    assert c["keys"].tolist() == K.KEYS
    assert c["flip"].tolist() == [1, 1]
    assert c["warm"].tolist() == ([1, 1] if case["cache"] is not None else [0, 0])
    want_lam = np.array([[float(x) for x in r] for r in case["lam"]])
    np.testing.assert_allclose(c["lam"], want_lam, rtol=1e-14, atol=1e-15)
    body = e.bodies()
    want_v = np.array([float(x) for x in case["v_b"]])
    np.testing.assert_allclose(body[0, 3:6], want_v, rtol=1e-14, atol=1e-15)
    # the static box keeps its (zero) velocity; B advanced by dt * v
    assert body[1, 3:6].tolist() == [0.0, 0.0, 0.0]
    np.testing.assert_allclose(body[0, :2], [1.0 + want_v[0] / 64.0, 3.0 + want_v[1] / 64.0],
                               rtol=1e-15)


@pytest.mark.parametrize("scene", ["random", "pile", "stacks"])
def test_reference_grid_broadphase_equals_brute_force(orc, scene):
    """The oracle's restatement of the reference's grid (Grid.hs:67-141,
    axes GridAxis 20 1 (-10), whose flattened cells alias across rows)
    gives exactly the brute-force culledKeys list (Aabb.hs:168-183), step
    after step, on scattered, piled and stacked worlds."""
    if scene == "random":
        w, cfg, steps = scenes.random_world(9, count=80, spacing=1.2), scenes.SceneConfig(), 40
    elif scene == "pile":
        sc = scenes.polygon_pile(count=1500, cols=40)
        w, cfg, steps = sc.worlds[0], sc.config, 30
    else:
        sc = scenes.reference_stacks()
        w, cfg, steps = sc.worlds[0], sc.config, 10
    a = orc.OracleEngine(w, cfg, broadphase=orc.BROADPHASE_GRID)
    b = orc.OracleEngine(w, cfg, broadphase=orc.BROADPHASE_BRUTE)
    for _ in range(steps):
        a.step(1)
        b.step(1)
        pa, pb = a.pairs(), b.pairs()
        assert pa.shape == pb.shape and (pa == pb).all()
    assert len(pa) > 0
    assert (a.bodies().view(np.uint64) == b.bodies().view(np.uint64)).all()


# ------------------------------------------------------------ circles and GJK

def test_circle_benchmark_contact(orc):
    """contact (Circle (-1,0) 1.1) (Circle (1,0) 1.1)
    (shapes/bench/Physics/Contact/Circle/Benchmark.hs:8-11), derived from
    Circle.hs:30-57: |ab| = 2, normal (1, 0) out of the first circle, the
    edges at x = 0.1 and -0.1 give the midpoint (0, 0), depth 2.2 - 2."""
    cs = orc.shape_contacts(("circle", 1.1, (-1.0, 0.0)), ("circle", 1.1, (1.0, 0.0)))
    assert cs == [(0, 0, 0, 1.0, 0.0, 0.0, 0.0, (1.1 + 1.1) - 2.0)]
    # apart by more than the radii: none; touching: depth 0
    assert orc.shape_contacts(("circle", 1.0, (-1.0, 0.0)), ("circle", 0.9, (1.0, 0.0))) == []
    t = orc.shape_contacts(("circle", 1.0, (-1.0, 0.0)), ("circle", 1.0, (1.0, 0.0)))
    assert t == [(0, 0, 0, 1.0, 0.0, 0.0, 0.0, 0.0)]


def test_gjk_benchmark_simplices(orc):
    """closestSimplex (makeOptBox 0 0 1.9 1.9) at (1,1), (0,1), (0,0.9)
    (shapes/bench/Physics/Contact/GJK/Benchmark.hs:21-28), derived by hand
    from GJK.hs:52-146: past the corner vertex 0 the search repeats vertex 0
    (a 1-simplex); above the top edge it extends to vertex 1 and stops on
    the edge (1, 0); inside, the 2-simplex (2, 0) extends to vertex 1 and
    the triangle (1, 2, 0) encloses the point."""
    b = box(0.0, 0.0, 1.9, 1.9)
    assert orc.closest_simplex(b, (1.0, 1.0)) == [0]
    assert orc.closest_simplex(b, (0.0, 1.0)) == [1, 0]
    assert orc.closest_simplex(b, (0.0, 0.9)) == [1, 2, 0]


def test_circle_hull_contacts_known_answers(orc):
    """CircleVsHull (CircleVsHull.hs:18-69) on the same box: a circle of
    radius 0.5 at (1,1) touches vertex 0 (normal from the centre towards the
    hull, depth 0.5 - sqrt(0.005)); at (0,1) the closest point of edge
    (1, 0) is (0, 0.95), depth 0.45, feature 1; a centre inside the hull
    gives no contact (the reference's 3-simplex case).  Circle as the second
    shape gives the same contact keyed (feature, 0) and Flip
    (Contact.hs:35-39)."""
    b = box(0.0, 0.0, 1.9, 1.9)
    (c,) = orc.shape_contacts(("circle", 0.5, (1.0, 1.0)), b)
    dx = 1.0 - 0.95  # the centre minus vertex 0, per axis
    d = math.sqrt(dx * dx + dx * dx)
    assert c[:3] == (0, 0, 0)
    assert c[3:7] == (-(dx / d), -(dx / d), 0.95, 0.95)
    assert c[7] == 0.5 - d
    (c,) = orc.shape_contacts(("circle", 0.5, (0.0, 1.0)), b)
    assert c[:3] == (0, 1, 0) and c[4:7] == (-1.0, 0.0, 0.95)
    assert abs(c[7] - 0.45) < 1e-15
    assert orc.shape_contacts(("circle", 0.5, (0.0, 0.9)), b) == []
    (f,) = orc.shape_contacts(b, ("circle", 0.5, (0.0, 1.0)))
    assert f[:3] == (1, 0, 1) and f[3:] == c[3:]


def _closest_on_hull(verts, o):
    """Closest point of a convex polygon's boundary to an outside point, by
    brute force over its edges."""
    best, bp = None, None
    n = len(verts)
    for k in range(n):
        a, b = np.array(verts[k]), np.array(verts[(k + 1) % n])
        t = np.clip(np.dot(np.array(o) - a, b - a) / np.dot(b - a, b - a), 0.0, 1.0)
        p = a + t * (b - a)
        dd = np.hypot(*(np.array(o) - p))
        if best is None or dd < best:
            best, bp = dd, p
    return best, bp


def test_circle_hull_closest_point_bruteforce(orc):
    """For points outside random convex polygons, GJK's contact point is the
    polygon's closest boundary point and the depth is r - distance, against
    a brute-force search over the edges (a wrong simplex shift or
    projection moves the point)."""
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(300):
        verts = [(x + 0.3, y - 0.2) for (x, y) in scenes.random_polygon(rng)]
        ang = rng.uniform(0, 2 * np.pi)
        o = (1.6 * np.cos(ang) + rng.uniform(-0.2, 0.2), 1.6 * np.sin(ang) + rng.uniform(-0.2, 0.2))
        dist, p = _closest_on_hull(verts, o)
        cs = orc.shape_contacts(("circle", 2.0, o), verts)
        assert len(cs) == 1
        (fa, fb, flip, nx, ny, px, py, depth) = cs[0]
        assert fa == 0 and flip == 0 and 0 <= fb < len(verts)
        assert np.hypot(px - p[0], py - p[1]) < 1e-9, (verts, o)
        assert abs(depth - (2.0 - dist)) < 1e-9
        assert abs(np.hypot(nx, ny) - 1.0) < 1e-12
        checked += 1
    assert checked == 300


def test_balls_scene_invariants(orc):
    """The reference's circle scene (Balls.makeScene, Balls.hs:68-71) on the
    oracle: every contact involving a circle carries a circle feature key 0,
    and the lagrangians stay in the friction cone with lambda_n >= 0."""
    sc = scenes.balls(dims=(6, 6))
    w = sc.worlds[0]
    e = orc.OracleEngine(w, sc.config)
    for _ in range(80):
        e.step(1)
        c = e.contacts()
        circ = w.radius > 0
        for (i, j, fi, fj), fl in zip(c["keys"].tolist(), c["flip"].tolist()):
            if circ[i]:
                assert fi == 0
            if circ[j]:
                assert fj == 0
            if circ[i] != circ[j]:
                assert fl == (0 if circ[i] else 1)
        lam = c["lam"]
        assert (lam[:, 0] >= 0).all()
    assert len(c["keys"]) > 30
