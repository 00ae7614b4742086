"""Oracle BVH builders and sorts: fixtures, closed forms and invariants.

* linked merge sort: the reference fixture of src/traversal_tests.rs:2824-2905
  (lengths 0..257, LCG keys in -6..6), expected = sort by (key, index);
* stable pair sort: numpy's stable argsort;
* Morton codes: closed-form corner/centre values and bit de-interleaving;
* Karras LBVH: equal to the textbook top-down construction that splits each
  range at the highest differing bit of (key, index);
* SAH: src/scene.rs:3627-3674 (250k triangles, depth <= 48, cost < 3) and
  src/scene.rs:4440-4466 (cover every triangle once, contain children).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def test_linked_sort_matches_stdlib_sort():
    for length in [0, 1, 2, 3, 5, 17, 63, 64, 129, 257]:
        nodes = np.zeros((max(1, length), 2), np.uint32)
        for i in range(length):
            k = ((((i * 1664525) + 1013904223) & 0xFFFFFFFF) >> 16) % 13 - 6
            nodes[i, 0] = np.uint32(k & 0xFFFFFFFF)
            nodes[i, 1] = 0xFFFFFFFF if i + 1 == length else i + 1
        expected = sorted(range(length), key=lambda j: (int(nodes[j, 0].view(np.int32)), j))
        head = O.linked_sort(nodes, 0 if length else 0xFFFFFFFF)
        order = []
        while head != 0xFFFFFFFF and len(order) <= length:
            order.append(head)
            head = int(nodes[head, 1])
        assert order == expected
        assert head == 0xFFFFFFFF


@pytest.mark.parametrize("n", [0, 1, 2, 7, 1000, 65537])
def test_stable_sort_pairs(n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 50, n).astype(np.uint32)
    perm = O.stable_sort_pairs(keys)
    assert np.array_equal(perm, np.argsort(keys, kind="stable"))


def deinterleave(code):
    out = [0, 0, 0]
    for bit in range(10):
        for axis in range(3):
            if code >> (3 * bit + (2 - axis)) & 1:
                out[axis] |= 1 << bit
    return out


def test_morton_closed_forms():
    lo = np.zeros((5, 4), np.float32)
    hi = np.zeros((5, 4), np.float32)
    pts = [(0, 0, 0), (8, 4, 2), (4, 2, 1), (8, 0, 0), (0, 4, 0)]
    for i, p in enumerate(pts):
        lo[i, :3] = p
        hi[i, :3] = p
    codes = O.morton_codes(lo, hi)
    assert codes[0] == 0
    assert codes[1] == 0x3FFFFFFF
    assert deinterleave(int(codes[2])) == [512, 512, 512]
    assert deinterleave(int(codes[3])) == [1023, 0, 0]
    assert deinterleave(int(codes[4])) == [0, 1023, 0]
    assert codes[3] == 0x24924924


def test_morton_degenerate_axis_is_zero():
    lo = np.zeros((3, 4), np.float32)
    lo[:, 0] = [0, 1, 2]
    codes = O.morton_codes(lo, lo.copy())
    assert [deinterleave(int(c)) for c in codes] == [[0, 0, 0], [512, 0, 0], [1023, 0, 0]]


def reference_hierarchy(keys, lo_i, hi_i, out):
    """Top-down: split [lo, hi] at the highest bit where (key, index) differ."""
    if lo_i == hi_i:
        return
    def full(i):
        return (int(keys[i]) << 32) | i
    a, b = full(lo_i), full(hi_i)
    prefix = 64 - (a ^ b).bit_length()
    split = lo_i
    step = hi_i - lo_i
    while step > 1:
        step = (step + 1) >> 1
        m = split + step
        if m < hi_i and 64 - (a ^ full(m)).bit_length() > prefix:
            split = m
    out[(lo_i, hi_i)] = split
    reference_hierarchy(keys, lo_i, split, out)
    reference_hierarchy(keys, split + 1, hi_i, out)


def node_ranges(nodes, leaf_max):
    """Walk the oracle's layout, returning {range: split} for internal nodes
    and the list of leaf ranges."""
    splits, leaves = {}, []
    def rng_of(i):
        n = nodes[i]
        if n[7] > 0:
            return (int(n[3]), int(n[3]) + int(n[7]) - 1)
        l = rng_of(int(n[3]))
        r = rng_of(int(n[3]) + 1)
        splits[(l[0], r[1])] = l[1]
        return (l[0], r[1])
    def walk(i):
        n = nodes[i]
        if n[7] > 0:
            leaves.append((int(n[3]), int(n[7])))
            return
        walk(int(n[3])); walk(int(n[3]) + 1)
    rng_of(0)
    walk(0)
    return splits, leaves


@pytest.mark.parametrize("n,dup", [(2, False), (5, False), (300, False), (300, True), (5000, True)])
def test_karras_matches_top_down_hierarchy(n, dup):
    rng = np.random.default_rng(n + dup)
    c = rng.random((n, 3)).astype(np.float32)
    if dup:
        c = np.round(c * 8) / 8
    lo = np.zeros((n, 4), np.float32); lo[:, :3] = c - 0.01
    hi = np.zeros((n, 4), np.float32); hi[:, :3] = c + 0.01
    codes, perm, nodes = O.lbvh(lo, hi, leaf_max=1)
    sk = codes[perm]
    expected = {}
    reference_hierarchy(sk, 0, n - 1, expected)
    splits, leaves = node_ranges(nodes, 1)
    assert splits == expected
    assert sorted(l[0] for l in leaves) == list(range(n))


def check_tree(nodes, lo, hi, perm, leaf_max):
    """Coverage once, exact leaf boxes, children contained (reachable nodes)."""
    n = lo.shape[0]
    visits = np.zeros(n, np.int64)
    stack = [(0, 0)]
    depth = 0
    f = nodes.view(np.float32)
    while stack:
        i, d = stack.pop()
        depth = max(depth, d)
        nd = nodes[i]
        if nd[7] > 0:
            assert nd[7] <= leaf_max
            idx = perm[nd[3]:nd[3] + nd[7]]
            visits[nd[3]:nd[3] + nd[7]] += 1
            assert np.all(lo[idx, :3] >= f[i, :3]) and np.all(hi[idx, :3] <= f[i, 4:7])
        else:
            for ch in (nd[3], nd[3] + 1):
                assert np.all(f[ch, :3] >= f[i, :3]) and np.all(f[ch, 4:7] <= f[i, 4:7])
                stack.append((int(ch), d + 1))
    assert np.all(visits == 1)
    return depth


def test_lbvh_invariants_random_triangles():
    rng = np.random.default_rng(3)
    n = 20000
    a = rng.random((n, 3)).astype(np.float32) * 10
    t = a[:, None, :] + rng.normal(size=(n, 3, 3)).astype(np.float32) * 0.05
    lo = np.zeros((n, 4), np.float32); lo[:, :3] = t.min(1)
    hi = np.zeros((n, 4), np.float32); hi[:, :3] = t.max(1)
    codes, perm, nodes = O.lbvh(lo, hi, leaf_max=4)
    assert check_tree(nodes, lo, hi, perm, 4) < 64
    # root box is the scene box, exactly
    f = nodes.view(np.float32)
    assert np.array_equal(f[0, :3], lo[:, :3].min(0)) and np.array_equal(f[0, 4:7], hi[:, :3].max(0))


def sah_fixture_250k():
    i = np.arange(250_000)
    x = np.where(i >= 249_900, np.float32(1000.0) + (i - 249_900).astype(np.float32) * np.float32(0.01),
                 (i % 500).astype(np.float32) * np.float32(0.01)).astype(np.float32)
    y = (i // 500).astype(np.float32) * np.float32(0.01)
    tri = np.zeros((i.size, 3, 3), np.float32)
    tri[:, 0, 0], tri[:, 0, 1] = x, y
    tri[:, 1, 0], tri[:, 1, 1] = x + np.float32(0.003), y
    tri[:, 2, 0], tri[:, 2, 1] = x, y + np.float32(0.003)
    lo = np.zeros((i.size, 4), np.float32); lo[:, :3] = tri.min(1)
    hi = np.zeros((i.size, 4), np.float32); hi[:, :3] = tri.max(1)
    return lo, hi


def test_sah_bounded_at_world_workload_size():
    lo, hi = sah_fixture_250k()
    perm, nodes = O.sah(lo, hi)
    depth = check_tree(nodes, lo, hi, perm, 250_000)
    assert depth <= 48
    assert O.tree_cost(nodes) < 3.0


def test_sah_covers_every_triangle_once_and_contains_children():
    n = 100
    i = np.arange(n, dtype=np.float32)
    lo = np.zeros((n, 4), np.float32); hi = np.zeros((n, 4), np.float32)
    lo[:, 0] = i; hi[:, 0] = i + 0.5; hi[:, 1] = 1.0
    perm, nodes = O.sah(lo, hi)
    check_tree(nodes, lo, hi, perm, n)
    perm0, nodes0 = O.sah(lo[:0], hi[:0])
    assert nodes0.shape[0] == 0


def test_lbvh_on_sah_fixture_is_valid():
    lo, hi = sah_fixture_250k()
    codes, perm, nodes = O.lbvh(lo, hi, leaf_max=4)
    check_tree(nodes, lo, hi, perm, 4)


def test_oracle_tlas_refit_is_a_tight_union():
    """refit (src/scene.rs:1040-1058): with the build's own instance boxes the
    TLAS is reproduced; after motion every reachable node is exactly the
    union of its children (internal) or instances (leaf), topology kept."""
    import dataclasses
    from bevy_penumbra_amd import scenes
    sc, motion = scenes.dynamic(seed=5, meshes=6, instances=40, stacks=6, slices=12)
    built = O.build_scene(sc)
    order = built.inst[:, 28].astype(np.uint32)
    roots = {int(r[28]): built.blas[int(r[25])] for r in built.inst}

    def boxes(fwd):
        lo = np.zeros((len(order), 4), np.float32)
        hi = np.zeros((len(order), 4), np.float32)
        for c in range(len(order)):
            lo[c], hi[c] = O.instance_bounds(fwd[c], roots[c])
        return lo, hi

    lo, hi = boxes(sc.inst_fwd)
    assert np.array_equal(O.tlas_refit(built.tlas, lo, hi, order), built.tlas)
    inv, fwd = scenes.dynamic_transforms(motion, 17)
    lo, hi = boxes(fwd)
    t = O.tlas_refit(built.tlas, lo, hi, order)
    f = t.view(np.float32)
    assert np.array_equal(t[:, [3, 7]], built.tlas[:, [3, 7]])       # first / count unchanged
    seen = 0
    stack = [0]
    while stack:
        k = stack.pop()
        seen += 1
        first, count = int(t[k, 3]), int(t[k, 7])
        if count:
            idx = order[first:first + count]
            assert np.array_equal(f[k, :3], lo[idx, :3].min(axis=0))
            assert np.array_equal(f[k, 4:7], hi[idx, :3].max(axis=0))
        else:
            assert np.array_equal(f[k, :3], np.minimum(f[first, :3], f[first + 1, :3]))
            assert np.array_equal(f[k, 4:7], np.maximum(f[first, 4:7], f[first + 1, 4:7]))
            stack += [first, first + 1]
    assert seen > 1


def test_sah_median_fallback_orders_by_total_cmp():
    """Coincident centroids take build_tree's median split, ordered by
    f32::total_cmp (src/scene.rs:2523-2533): -0 sorts before +0 (which `<`
    calls equal), ties keep the index order."""
    n = 8
    lo = np.zeros((n, 4), np.float32)
    hi = np.zeros((n, 4), np.float32)
    neg = [1, 4, 6]                                        # boxes at x = -0
    lo[neg, 0] = hi[neg, 0] = np.float32(-0.0)
    perm, nodes = O.sah(lo, hi)
    assert nodes.shape[0] == 3 and nodes[0, 7] == 0 and nodes[1, 7] == 4 and nodes[2, 7] == 4
    assert perm.tolist() == [1, 4, 6, 0, 2, 3, 5, 7]


def _model_build_tree(lo, hi, ties_by_index=False):
    """build_tree (src/scene.rs:2398-2554) in plain Python loops over boxes:
    12 bins with an f64 bin index, bin cost area(f64) x count, Hoare
    partition, leaves <= 4, depth 48, and for coincident centroids the median
    of select_nth_unstable_by(total_cmp), which for slices of at most 10
    elements is Rust's stable insertion sort (insertion_sort_shift_left).
    ties_by_index=True models the mistake of breaking key ties by index.
    Returns (perm, nodes as (first, count, min3, max3))."""
    f32 = np.float32
    n = lo.shape[0]
    perm = list(range(n))
    nodes = [None]

    def centre(e, a):
        return f32(f32(lo[e, a]) * f32(0.5) + f32(hi[e, a]) * f32(0.5))

    def key(e, a):
        u = int(np.array([centre(e, a)], np.float32).view(np.uint32)[0])
        return (~u & 0xFFFFFFFF) if u & 0x80000000 else (u | 0x80000000)

    def area(mn, mx):
        sx, sy, sz = (float(mx[a]) - float(mn[a]) for a in range(3))
        return 2.0 * (sx * sy + sy * sz + sz * sx)

    def vmin(vals):  # glam Vec3::min per component: a < b ? a : b, folded in order
        acc = math.inf
        for v in vals:
            acc = acc if acc < v else v
        return acc

    def vmax(vals):
        acc = -math.inf
        for v in vals:
            acc = acc if acc > v else v
        return acc

    def build(index, first, count, depth):
        el = perm[first:first + count]
        mn = [vmin(float(lo[e, a]) for e in el) for a in range(3)]
        mx = [vmax(float(hi[e, a]) for e in el) for a in range(3)]
        nodes[index] = (first, count, mn, mx)
        if count <= 4 or depth == 48:
            return
        cmin = [vmin(float(centre(e, a)) for e in el) for a in range(3)]
        cmax = [vmax(float(centre(e, a)) for e in el) for a in range(3)]
        best, split = math.inf, None
        for axis in range(3):
            span = cmax[axis] - cmin[axis]
            if span <= 0.0:
                continue
            binof = lambda c: min(int((float(c) - cmin[axis]) / span * 12), 11)  # noqa: E731
            bins = [[[math.inf] * 3, [-math.inf] * 3, 0] for _ in range(12)]
            for e in el:
                b = bins[binof(centre(e, axis))]
                b[0] = [min(b[0][a], float(lo[e, a])) for a in range(3)]
                b[1] = [max(b[1][a], float(hi[e, a])) for a in range(3)]
                b[2] += 1

            def join(x, y):
                return [[min(x[0][a], y[0][a]) for a in range(3)],
                        [max(x[1][a], y[1][a]) for a in range(3)], x[2] + y[2]]

            def cost(b):
                return 0.0 if b[2] == 0 else area(b[0], b[1]) * b[2]
            right = [None] * 12
            acc = [[math.inf] * 3, [-math.inf] * 3, 0]
            for i in range(11, -1, -1):
                acc = join(acc, bins[i])
                right[i] = acc
            left = [[math.inf] * 3, [-math.inf] * 3, 0]
            for i in range(11):
                left = join(left, bins[i])
                c = cost(left) + cost(right[i + 1])
                if left[2] > 0 and right[i + 1][2] > 0 and c < best:
                    best, split = c, (axis, i, span)
        if split is not None:
            axis, last, span = split
            l, r = 0, count
            while l < r:
                if min(int((float(centre(el[l], axis)) - cmin[axis]) / span * 12), 11) <= last:
                    l += 1
                else:
                    r -= 1
                    el[l], el[r] = el[r], el[l]
            mid = l
        else:
            size = [cmax[a] - cmin[a] for a in range(3)]
            axis = 0 if size[0] >= size[1] and size[0] >= size[2] else (1 if size[1] >= size[2] else 2)
            assert count <= 10, "quickselect of larger slices is not modelled"
            if ties_by_index:
                el.sort(key=lambda e: (key(e, axis), e))
            else:
                for i in range(1, count):          # insertion_sort_shift_left
                    j = i
                    while j > 0 and key(el[j], axis) < key(el[j - 1], axis):
                        el[j], el[j - 1] = el[j - 1], el[j]
                        j -= 1
            mid = count // 2
        perm[first:first + count] = el
        children = len(nodes)
        nodes.extend([None, None])
        nodes[index] = (children, 0, mn, mx)
        build(children, first, mid, depth + 1)
        build(children + 1, first + mid, count - mid, depth + 1)

    build(0, 0, n, 0)
    return perm, nodes


def coincident_centroid_boxes(group):
    """`group` boxes of different extents sharing the centre (1, 1, 1),
    interleaved in input order with boxes spread along x = 10..15, so the root's
    partition permutes the group before its median split."""
    spread = 6
    lo, hi = [], []
    for i in range(max(group, spread)):
        if i < group:
            s = np.array([0.0625 * (i + 1), 0.125 * ((i * 5) % 7 + 1), 0.03125 * ((i * 3) % 5 + 1)])
            lo.append(1.0 - s), hi.append(1.0 + s)
        if i < spread:
            c = np.array([10.0 + i, 2.0, 3.0])
            lo.append(c - 0.25), hi.append(c + 0.25)
    lo4 = np.zeros((len(lo), 4), np.float32)
    hi4 = np.zeros((len(hi), 4), np.float32)
    lo4[:, :3], hi4[:, :3] = lo, hi
    return lo4, hi4


@pytest.mark.parametrize("group", [5, 6, 7, 8, 9, 10])
def test_sah_median_keeps_partition_order_of_coincident_centroids(group):
    """The median split of 5-10 coincident centroids keeps the order a
    parent's partition left (stable insertion sort), which gives other split
    sets and node boxes than ordering the ties by index."""
    lo, hi = coincident_centroid_boxes(group)
    want_perm, want_nodes = _model_build_tree(lo, hi)
    wrong_perm, wrong_nodes = _model_build_tree(lo, hi, ties_by_index=True)
    assert want_perm != wrong_perm, "the fixture must reach a permuted median slice"
    perm, nodes = O.sah(lo, hi)
    assert perm.tolist() == want_perm
    assert nodes.shape[0] == len(want_nodes)
    f = nodes.view(np.float32)
    for k, (first, count, mn, mx) in enumerate(want_nodes):
        assert int(nodes[k, 3]) == first and int(nodes[k, 7]) == count, k
        assert np.array_equal(f[k, :3], np.float32(mn)) and np.array_equal(f[k, 4:7], np.float32(mx)), k
    if group >= 6:  # a median over >= 6 boxes: the wrong order moves a box to the other side
        assert any(a[2:] != b[2:] for a, b in zip(want_nodes, wrong_nodes)), "node boxes must tell the orders apart"


def test_sah_box_keeps_the_later_of_equal_zeros():
    """Node boxes fold glam's Vec3::min/max (a < b ? a : b): of +0 and -0 the
    later element's zero is kept -- in the oracle and in the plain-loop model;
    a fold that keeps the first one (fminf on some libms, Python's min) fails."""
    lo = np.zeros((2, 4), np.float32)
    hi = np.ones((2, 4), np.float32)
    hi[:, 3] = lo[:, 3] = 0
    for first, second in ((0.0, -0.0), (-0.0, 0.0)):
        lo[0, 0], lo[1, 0] = first, second
        hi[0, 1], hi[1, 1] = first, second                   # max over y: both zeros
        lo[:, 1] = -1.0
        perm, nodes = O.sah(lo, hi)
        assert perm.tolist() == [0, 1]
        mn_x = nodes[0, 0:1].view(np.float32)[0]
        mx_y = nodes[0, 5:6].view(np.float32)[0]
        assert mn_x == 0.0 and np.signbit(mn_x) == np.signbit(np.float32(second))
        assert mx_y == 0.0 and np.signbit(mx_y) == np.signbit(np.float32(second))
        _, mnodes = _model_build_tree(lo, hi)
        assert np.signbit(np.float32(mnodes[0][2][0])) == np.signbit(np.float32(second))
        assert np.signbit(np.float32(mnodes[0][3][1])) == np.signbit(np.float32(second))
