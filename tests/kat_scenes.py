"""Hand-derived known answers for the contact update, shared by the oracle
tests (CPU) and the drop-in GPU tests.

The scene is the reference's contact-benchmark pair
(shapes/bench/Physics/Contact/Benchmark.hs:16-27): body 0 is the 2x2 box B
at (1, 3), dynamic (m = 1, I = 1/2, so inverse masses (1, 1, 2)); body 1 is
the 4x4 box A at the origin, static.  They touch along y = 2, so the pair
(1, 0) has two Flip contacts on B's edge 2 (normal (0, -1)), keys
((1,0),(1,2)) at p = (0, 2) and ((1,0),(0,2)) at p = (2, 2), depth 0 (no
Baumgarte bias).  The rows, derived from NonPenetration.hs:34-44 and
Friction.hs:31-44 with the Flip swap (Utils.hs:168-209): B's normal row is
(0, 1, -1) at p = (0, 2) and (0, 1, 1) at p = (2, 2); its friction row is
(1, 0, 1) at both.  Gravity is 0, so each case is the warm-start pass and
the sweeps of one step.
"""
from fractions import Fraction as F

from shapes_amd import scenes

KEYS = [[1, 0, 1, 2], [1, 0, 0, 2]]


def box(w, h):
    return scenes.rectangle_vertices(w, h)


def world(v_b=(0.0, 0.0), w_b=0.0, mu=0.5, bounce_b=0.0, bounce_a=0.0):
    b = scenes._Builder()
    b.add(box(2.0, 2.0), (1.0, 3.0), vel=v_b, omega=w_b, mass=1.0, inertia=0.5, mu=mu,
          bounce=bounce_b)
    b.add(box(4.0, 4.0), (0.0, 0.0), mu=mu, bounce=bounce_a)
    return b.build()


def config(iterations):
    return scenes.SceneConfig(dt=1.0 / 64.0, gy=0.0, iterations=iterations)


# (1) useCache (Solvers/Contact.hs:100-114): both cached solutions applied in
# contact order, normal then friction, no sweep.  v_B = sum of (J * lam) * im:
# (1, 0.5) at p = (0,2): (0, 1, -2) + (0.5, 0, 1); (2, -0.25) at p = (2,2):
# (0, 2, 4) + (-0.25, 0, -0.5).  Exact in binary.
WARM = {
    "world": dict(),
    "iterations": 0,
    "cache": [[1.0, 0.5], [2.0, -0.25]],
    "v_b": [F(1, 4), F(3), F(5, 2)],
    "lam": [[F(1), F(1, 2)], [F(2), F(-1, 4)]],
}

# (2) One sweep from rest contact with v_B = (0, -2), e = min(1, 1/2) = 1/2
# (bounceB, Restitution.hs:34-47, closing velocity with the unswapped radii
# and the normal negated for Flip: (0, 1)), mu = 1/2.
#   p = (0,2): Jn.v = -2, bounce = 1/2 * -2 = -1, mn = 3 -> ln = 1; Jt.v = 0
#     -> lt = 0; v_B = (0, -1, -2).
#   p = (2,2): closing = (0, -1) + zcross(-2, (1, -1)) = (-2, -3) -> bounce
#     -3/2; Jn.v = -3 -> ln = 3/2; friction from the SAME pre-update v:
#     Jt.v = -2 -> lt = 2/3 (bound 3/2 * 1/2 = 3/4); v_B += (0, 3/2, 3), then
#     += (2/3, 0, 4/3) -> (2/3, 1/2, 7/3).
BOUNCE = {
    "world": dict(v_b=(0.0, -2.0), bounce_b=0.5, bounce_a=1.0),
    "iterations": 1,
    "cache": None,
    "v_b": [F(2, 3), F(1, 2), F(7, 3)],
    "lam": [[F(1), F(0)], [F(3, 2), F(2, 3)]],
}

# (3) The same with mu = 1/10: the second contact's friction is clamped at
# its accumulated normal impulse times mu (3/2 * 1/10 = 3/20).
BOUNCE_CLAMPED = {
    "world": dict(v_b=(0.0, -2.0), bounce_b=0.5, bounce_a=1.0, mu=0.1),
    "iterations": 1,
    "cache": None,
    "v_b": [F(3, 20), F(1, 2), F(13, 10)],
    "lam": [[F(1), F(0)], [F(3, 2), F(3, 20)]],
}

# (4) Two sweeps of (2).  Sweep 2, p = (0,2) (cached (1, 0)): closing =
# (2/3, 1/2) + zcross(7/3, (-1, -1)) = (3, -11/6), bounce -11/12, Jn.v =
# -11/6 -> new ln = 11/12, accumulated 23/12; Jt.v = 3 -> new lt = -1,
# clamped to -(23/12 * 1/2) = -23/24 (the bound uses the accumulated normal
# impulse, Constraints/Contact.hs:106-109); v_B = (-7/24, 17/12, -17/12).
# p = (2,2) (cached (3/2, 2/3)): closing (-41/24, 0), bounce 0, Jn.v = 0 ->
# new ln = -0, accumulated 3/2; Jt.v = -41/24 -> accumulated 2/3 + 41/72 =
# 89/72, clamped to 3/4, applied 1/12; v_B = (-5/24, 17/12, -5/4).
BOUNCE_TWO_SWEEPS = {
    "world": dict(v_b=(0.0, -2.0), bounce_b=0.5, bounce_a=1.0),
    "iterations": 2,
    "cache": None,
    "v_b": [F(-5, 24), F(17, 12), F(-5, 4)],
    "lam": [[F(23, 12), F(-23, 24)], [F(3, 2), F(3, 4)]],
}

CASES = {"warm": WARM, "bounce": BOUNCE, "bounce_clamped": BOUNCE_CLAMPED,
         "bounce_two_sweeps": BOUNCE_TWO_SWEEPS}
