"""Scene generators: seeded, convex CCW hulls centred on the centroid,
closed-form mass properties (CPU)."""
import numpy as np

from shapes_amd import scenes


def _check_world(w):
    assert w.hull_off[0] == 0 and len(w.hull_off) == w.n + 1
    for i in range(w.n):
        xs = w.hull_x[w.hull_off[i]:w.hull_off[i + 1]]
        ys = w.hull_y[w.hull_off[i]:w.hull_off[i + 1]]
        assert 3 <= len(xs) <= 8
        x1, y1 = np.roll(xs, -1), np.roll(ys, -1)
        # every turn is to the left: convex and counter-clockwise
        ex, ey = x1 - xs, y1 - ys
        turn = ex * np.roll(ey, -1) - ey * np.roll(ex, -1)
        assert (turn > 0).all()
        if w.mass[i] > 0:
            m, (cx, cy), inertia = scenes.polygon_mass_properties(xs, ys)
            assert abs(m - w.mass[i]) < 1e-12 * max(1, m)
            assert abs(cx) < 1e-12 and abs(cy) < 1e-12
            assert abs(inertia - w.inertia[i]) < 1e-12 * max(1, inertia)


def test_polygon_pile_and_pyramid_are_valid():
    _check_world(scenes.pyramid().worlds[0])
    _check_world(scenes.polygon_pile(count=300).worlds[0])
    _check_world(scenes.random_world(3))


def test_vectorized_batches_are_valid_and_seeded():
    a = scenes.world_batch_fast(n_worlds=12)
    b = scenes.world_batch_fast(n_worlds=12)
    for wa, wb in zip(a.worlds, b.worlds):
        _check_world(wa)
        assert np.array_equal(wa.px, wb.px) and np.array_equal(wa.hull_x, wb.hull_x)
    sizes = [w.n - 1 for w in a.worlds]
    assert min(sizes) >= 8 and max(sizes) <= 1024
    p = scenes.pile_batch(n_worlds=3)
    for w in p.worlds:
        _check_world(w)
        assert 2048 <= w.n - 1 <= 32768


def test_reference_stacks_layout():
    """Stacks.makeScene (30,30) 0: floor + 900 boxes, columns left to right
    (x from -2.9 in steps of 0.2), bottom to top from y = -4.5."""
    w = scenes.reference_stacks().worlds[0]
    assert w.n == 901
    assert w.px[1] == 0.0 - (0.2 * 29.0 / 2.0) and w.py[1] == -4.5
    assert w.px[31] == w.px[1] + 0.2 and w.py[2] == -4.5 + 0.2
