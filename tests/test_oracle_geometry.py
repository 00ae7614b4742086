"""Oracle ray/box/triangle primitives against the reference's fixtures.

Shared-edge fixture: src/traversal_tests.rs:2909-2998 (scales 1e-12..1e12,
both ray signs, upper and lower triangle of a split square; the ray meets the
shared diagonal exactly, so both triangles must report it).
"""
import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("scale", [1e-12, 1e-6, 1.0, 1e6, 1e12])
@pytest.mark.parametrize("sign", [-1.0, 1.0])
@pytest.mark.parametrize("upper", [False, True])
def test_shared_edges_are_watertight_and_scale_aware(scale, sign, upper):
    s = np.float32(scale)
    a = np.array([-s, -s, 0], np.float32)
    diag = np.array([s, s, 0], np.float32)
    b = diag if upper else np.array([s, -s, 0], np.float32)
    c = np.array([-s, s, 0], np.float32) if upper else diag
    origin = np.array([0, 0, s * np.float32(sign)], np.float32)
    direction = np.array([0, 0, -sign], np.float32)
    shorter = s * (np.float32(1.0 if sign < 0 else 2.0) if upper else np.float32(0.5))
    t, w1, w2 = O.intersect_triangle(origin, direction, a, b, c)
    assert abs(t / s - 1.0) < 1e-5
    assert w1 >= 0 and w2 >= 0 and w1 + w2 <= 1
    mn = np.minimum(np.minimum(a, b), c)
    mx = np.maximum(np.maximum(a, b), c)
    initial, entry, _ = O.box_interval(origin, direction, mn, mx, s * 2)
    updated, _, _ = O.box_interval(origin, direction, mn, mx, shorter)
    cached = initial and entry <= shorter
    checks = int(initial) | (int(updated) << 1) | (int(cached) << 2)
    assert checks == (7 if upper else 1)


def test_box_interval_zero_direction_components():
    o = np.array([0.5, 0.5, -3], np.float32)
    d = np.array([0, 0, 1], np.float32)
    hit, e, x = O.box_interval(o, d, np.zeros(3, np.float32), np.ones(3, np.float32), 100)
    assert hit and e == 3.0 and x >= 4.0
    o2 = np.array([1.5, 0.5, -3], np.float32)
    hit, e, x = O.box_interval(o2, d, np.zeros(3, np.float32), np.ones(3, np.float32), 100)
    assert not hit and e == 0.0 and x == 0.0


def test_box_far_endpoint_is_conservative():
    # exit of 1e6 padded by |exit| * 3.5762791e-7 (traversal.slang:364)
    o = np.zeros(3, np.float32)
    d = np.array([1, 0, 0], np.float32)
    hit, e, x = O.box_interval(o, d, np.array([1e6 - 1, -1, -1], np.float32),
                               np.array([1e6, 1, 1], np.float32), 1e7)
    assert hit and x > 1e6 and x == np.float32(1e6) + np.float32(np.float32(1e6) * np.float32(3.5762791e-7))


def test_triangle_miss_and_degenerate():
    o = np.zeros(3, np.float32)
    d = np.array([0, 0, 1], np.float32)
    a, b, c = (np.array(v, np.float32) for v in ([1, 1, 5], [2, 1, 5], [1, 2, 5]))
    assert O.intersect_triangle(o, d, a, b, c)[0] == -1
    # degenerate (collinear) triangle through the ray
    a, b, c = (np.array(v, np.float32) for v in ([-1, 0, 5], [0, 0, 5], [1, 0, 5]))
    assert O.intersect_triangle(o, d, a, b, c)[0] == -1
    # zero direction is invalid
    a, b, c = (np.array(v, np.float32) for v in ([-1, -1, 5], [1, -1, 5], [0, 1, 5]))
    assert O.intersect_triangle(o, np.zeros(3, np.float32), a, b, c)[0] == -1
    t, w1, w2 = O.intersect_triangle(o, d, a, b, c)
    assert t == 5.0
    # barycentric of the ray point (0,0): analytic weights for b and c
    assert abs(w1 - 0.25) < 1e-7 and abs(w2 - 0.5) < 1e-7


def _edge_reference(ax, ay, bx, by):
    """orientedEdge as the reference writes it (traversal.slang:379-392), in
    float32: order the points, edge of the ordered pair, negate if swapped."""
    f32 = np.float32
    reverse = (ax > bx) | ((ax == bx) & (ay > by))
    fx, fy = np.where(reverse, bx, ax), np.where(reverse, by, ay)
    sx, sy = np.where(reverse, ax, bx), np.where(reverse, ay, by)
    pa, pb = fx * sy, fy * sx
    edge = pa - pb
    err = (np.abs(pa) + np.abs(pb)) * f32(2.3841858e-7)
    edge = np.where(np.abs(edge) <= err, f32(0.0), edge)
    return np.where(reverse, -edge, edge).astype(f32)


def _edge_unordered(ax, ay, bx, by):
    """The kernels' form: a.x b.y - a.y b.x, a zeroed edge signed by the swap."""
    f32 = np.float32
    reverse = (ax > bx) | ((ax == bx) & (ay > by))
    pa, pb = ax * by, ay * bx
    edge = pa - pb
    err = (np.abs(pa) + np.abs(pb)) * f32(2.3841858e-7)
    return np.where(np.abs(edge) <= err, np.where(reverse, f32(-0.0), f32(0.0)), edge).astype(f32)


def test_oriented_edge_needs_no_point_ordering():
    """The swapped products are the same products and rounding is symmetric,
    so the kernels' unordered edge equals the reference's ordered one bit for
    bit, zero signs included (the barycentric outputs divide by them)."""
    rng = np.random.default_rng(11)
    n = 1 << 20
    v = (rng.uniform(-2, 2, (4, n)) * np.exp2(rng.integers(-40, 1, (4, n)))).astype(np.float32)
    # ties on x, equal points, zeros and collinear pairs
    v[2, : n // 8] = v[0, : n // 8]
    v[3, : n // 16] = v[1, : n // 16]
    v[:, n // 8: n // 8 + 4096] = 0.0
    k = np.float32(rng.uniform(-3, 3))
    v[2, n // 4: n // 2], v[3, n // 4: n // 2] = v[0, n // 4: n // 2] * k, v[1, n // 4: n // 2] * k
    with np.errstate(all="ignore"):
        a = _edge_reference(*v)
        b = _edge_unordered(*v)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert (a == 0).sum() > n // 8 and (np.signbit(a) & (a == 0)).sum() > 0
