"""Oracle of the temporal interval history (mergeEntry's history part).

Pinned on the reference's own temporal cases (temporalCases,
src/split_merge_tests.rs:31-76, expectations :1303-1359): 512 observations
saturate both masses at 1 / (1 - blend); one opposite observation moves J to
2 j and beta to 1 - b; 200 frames without observations keep value and mass
bit for bit; four observations at once retain j^4 and b^4; each channel's
history is used only when its blend is positive.  Beyond those: the masses
follow the geometric series, J is the geometric-weighted mean of the
observations, a dirty box drops exactly the intervals whose retained
dependency meets it (hand-built dependencies: bounded, unknown, measured
empty; a reused value's dependency is the union of the old and the new), the
trace's dependencies are its segments clipped to the scene's root box, and the
merge of a first update with history is the plain merge.
"""
import numpy as np
import pytest

from oracle import oracle as O


def known_empty(n):
    """Dependencies of n intervals: known, measured empty."""
    d = np.zeros((n, 8), np.float32)
    d[:, 3] = 1.0
    return d


def one_interval():
    """A grid with a single interval (one probe, one direction, one level)."""
    return O.gi_params((0.0, 0.0, 0.0), 1.0, (1, 1, 1), 1, base_phi=1, base_y=1)


class Hist:
    """History of the one interval, driven like the reference's advance()."""

    def __init__(self, rb, tb, frames=30, persistent=1):
        self.p = one_interval()
        self.rb, self.tb, self.frames, self.persistent = rb, tb, frames, persistent
        self.h = np.zeros((1, 16), np.float32)
        self.valid = False

    def advance(self, radiance, beta, count):
        iv = np.array([[radiance] * 3 + [beta if count else -1.0]], np.float32)
        out, self.h = O.gi_history(self.p, self.rb, self.tb, self.frames, self.persistent, self.valid, iv,
                                   np.array([count], np.uint32), self.h, known_empty(1))
        self.valid = True
        return out[0]

    @property
    def masses(self):
        return float(self.h[0, 3]), float(self.h[0, 5])


def close(a, b):
    assert abs(a - b) < 2e-4, (a, b)


@pytest.mark.parametrize("j,beta", [(0.5, 0.75), (0.9, 0.95)])
def test_reference_temporal_cases(j, beta):
    h = Hist(j, beta)
    for _ in range(512):
        h.advance(2.0, 0.0, 1)
    close(h.masses[0], 1.0 / (1.0 - j))
    close(h.masses[1], 1.0 / (1.0 - beta))
    after = h.advance(0.0, 1.0, 1)
    close(after[0], 2.0 * j)
    close(after[3], 1.0 - beta)
    state = h.h.copy()
    for _ in range(200):
        kept = h.advance(0.0, 1.0, 0)
    assert np.array_equal(kept.view(np.uint32), after.view(np.uint32))    # no decay of world state
    assert np.array_equal(h.h[0, :6].view(np.uint32), state[0, :6].view(np.uint32))  # nor of mass
    for _ in range(512):
        h.advance(2.0, 0.0, 1)
    four = h.advance(0.0, 1.0, 4)
    close(four[0], 2.0 * j ** 4)
    close(four[3], 1.0 - beta ** 4)
    close(h.masses[0], 1.0 / (1.0 - j))
    close(h.masses[1], 1.0 / (1.0 - beta))


def test_reference_single_channel_histories():
    beta_only = Hist(0.0, 0.9)
    for _ in range(512):
        beta_only.advance(2.0, 0.0, 1)
    v = beta_only.advance(0.0, 1.0, 1)
    close(v[0], 0.0)
    close(v[3], 0.1)
    for _ in range(200):
        beta_only.advance(0.0, 1.0, 0)
    close(beta_only.masses[0], 0.0)
    close(beta_only.masses[1], 10.0)
    radiance_only = Hist(0.9, 0.0)
    for _ in range(512):
        radiance_only.advance(2.0, 0.0, 1)
    v = radiance_only.advance(0.0, 1.0, 1)
    close(v[0], 1.8)
    close(v[3], 1.0)
    for _ in range(200):
        radiance_only.advance(0.0, 1.0, 0)
    close(radiance_only.masses[0], 10.0)
    close(radiance_only.masses[1], 0.0)


@pytest.mark.parametrize("b", [0.3, 0.88, 0.98])
def test_masses_follow_the_geometric_series_and_j_the_weighted_mean(b):
    rng = np.random.default_rng(7)
    xs = rng.uniform(0.0, 5.0, 12)
    h = Hist(b, b)
    for k, x in enumerate(xs, 1):
        v = h.advance(float(x), 0.25, 1)
        assert abs(h.masses[0] - (1 - b ** k) / (1 - b)) <= 2e-6 * k / (1 - b)
        w = b ** np.arange(k - 1, -1, -1)
        want = float((xs[:k] * w).sum() / w.sum())
        assert abs(v[0] - want) <= 1e-5 * max(1.0, want), (k, v[0], want)
        close(v[3], 0.25)


def test_history_off_or_expired_starts_over():
    h = Hist(0.9, 0.9, frames=0)
    h.advance(2.0, 0.0, 1)
    v = h.advance(4.0, 0.5, 1)
    assert v[0] == 4.0 and v[3] == 0.5                    # history_frames 0: nothing reused
    e = Hist(0.9, 0.9, frames=3, persistent=0)
    e.p.frame = 10
    e.advance(2.0, 0.0, 1)
    e.p.frame = 13
    close(e.advance(4.0, 0.0, 1)[0], (4.0 * 1.0 + 2.0 * 0.9) / 1.9)   # 3 frames old: alive
    e.p.frame = 20
    assert e.advance(4.0, 0.0, 0)[3] == 1.0               # 7 frames since the last sample: expired


def test_dirty_box_drops_exactly_the_intervals_whose_dependency_it_meets():
    """Hand-built dependencies of eight intervals: boxes [i + 0.1, i + 0.9]
    (x) x [0.1, 0.9]^2 for i = 0..4, an unknown one, a measured-empty one and
    a box over x in [3, 6].  A dirty box over x in [3.95, 4.05] lies between
    the boxes of intervals 3 and 4: only the unknown dependency and the box
    over [3, 6] are dropped.  A second dirty box over [4.5, 4.6] drops
    interval 4 as well.  Without dirty boxes everything is reused, the unknown
    dependency included.  A box touching a dependency at one face meets it."""
    p = O.gi_params((0.0, 0.0, 0.0), 1.0, (8, 1, 1), 1, base_phi=1, base_y=1, terminal_tail=0)
    n = O.total_intervals(p)
    assert n == 8
    dep = np.zeros((n, 8), np.float32)
    for i in range(5):
        dep[i] = [i + 0.1, 0.1, 0.1, 1.0, i + 0.9, 0.9, 0.9, 1.0]
    dep[5] = 0.0                                                    # unknown
    dep[6] = [0, 0, 0, 1.0, 0, 0, 0, 0.0]                           # measured empty
    dep[7] = [3.0, 0.2, 0.2, 1.0, 6.0, 0.8, 0.8, 1.0]
    iv = np.zeros((n, 4), np.float32)
    iv[:, 0], iv[:, 3] = 3.0, 0.5
    ns = np.ones(n, np.uint32)
    _, h1 = O.gi_history(p, 0.9, 0.9, 30, 1, False, iv, ns, np.zeros((n, 16), np.float32), dep)
    assert np.array_equal(h1[:, 8:16], dep)                         # the first pass stores this trace's
    iv2 = iv.copy()
    iv2[:, 0] = 1.0
    blended = (1.0 + 3.0 * 0.9) / 1.9
    box = ((3.95, 0.0, 0.0), (4.05, 1.0, 1.0))
    out, h2 = O.gi_history(p, 0.9, 0.9, 30, 1, True, iv2, ns, h1, known_empty(n), boxes=[box])
    fresh = out[:, 0] == 1.0
    assert fresh.tolist() == [False, False, False, False, False, True, False, True]
    assert abs(out[0, 0] - blended) < 1e-6
    out, _ = O.gi_history(p, 0.9, 0.9, 30, 1, True, iv2, ns, h1, known_empty(n),
                          boxes=[box, ((4.5, 0.4, 0.4), (4.6, 0.6, 0.6))])
    assert (out[:, 0] == 1.0).tolist() == [False, False, False, False, True, True, False, True]
    out, _ = O.gi_history(p, 0.9, 0.9, 30, 1, True, iv2, ns, h1, known_empty(n))
    assert np.all(out[:, 0] != 1.0)                                 # no dirty box: all reused
    out, _ = O.gi_history(p, 0.9, 0.9, 30, 1, True, iv2, ns, h1, known_empty(n),
                          boxes=[((2.9, 0.0, 0.0), (2.9, 0.5, 0.5))])
    assert (out[:, 0] == 1.0).tolist() == [False, False, True, False, False, True, False, False]
    out, _ = O.gi_history(p, 0.9, 0.9, 30, 1, True, iv2, ns, h1, known_empty(n), root_grew=True)
    assert (out[:, 0] == 1.0).tolist() == [False] * 6 + [True, False]   # a grown root: empty is stale
    # a reused value keeps the old dependency joined with the new; a dropped
    # one takes the new alone
    new = np.zeros((n, 8), np.float32)
    new[:] = [10.0, 0.0, 0.0, 1.0, 11.0, 1.0, 1.0, 1.0]
    _, h3 = O.gi_history(p, 0.9, 0.9, 30, 1, True, iv2, ns, h1, new, boxes=[box])
    assert np.array_equal(h3[0, 8:16], np.float32([0.1, 0.0, 0.0, 1.0, 11.0, 1.0, 1.0, 1.0]))
    assert np.array_equal(h3[5, 8:16], new[5]) and np.array_equal(h3[7, 8:16], new[7])
    assert np.array_equal(h3[6, 8:16], new[6])                      # empty joined with new = new


def test_trace_dependencies_are_the_segments_clipped_to_the_root():
    """The dense fixture (boxes in [-3.9, 2.0] x [-0.9, 3.04] x [-2.0, 4.03]):
    every interval that missed (beta = 1) depends on its own segment from
    c + d * start, clipped to the root box and to the interval's length,
    widened by max(scale * 1e-6, 2^-23) -- checked in f64 with a tolerance;
    a segment that misses the root is measured empty; an interval that hit
    has a bounded dependency."""
    from tests.transport_cases import dense_fixture_params, dense_fixture_scene
    sc = dense_fixture_scene()
    built = O.build_scene(sc)
    p = dense_fixture_params()
    iv, ns, dep = O.gi_trace(built, O.build_lights(sc), sc.materials, p, threads=8, dep=True)
    assert np.all(dep[:, 3] == 1.0)                                  # every dependency known
    root = built.tlas[0].view(np.float32)
    lo, hi = root[:3].astype(np.float64), root[4:7].astype(np.float64)
    checked = empty = hits = 0
    for L in range(p.levels):
        a, b = O.level_offset(p, L), O.level_offset(p, L + 1)
        start, end = O.interval_bounds(p, L)
        nd, dims = O.level_dims(p, L)
        D = O.level_directions(p, L)
        rng = np.random.default_rng(L)
        for g in rng.choice(np.arange(a, b), size=400, replace=False):
            local = g - a
            probe, k = divmod(local, D)
            ix, iy, iz = probe % dims[0], (probe // dims[0]) % dims[1], probe // (dims[0] * dims[1])
            sl = 2.0 ** L
            c = np.array([-8.0 + (ix + 0.5) * sl, -8.0 + (iy + 0.5) * sl, -8.0 + (iz + 0.5) * sl])
            from tests.transport_cases import _dense_offset
            d = O.bin_direction(k, p.base_phi << L, p.base_y << L, *_dense_offset(p.seed)).astype(np.float64)
            o = c + d * start
            with np.errstate(divide="ignore", invalid="ignore"):
                t0, t1 = (lo - o) / d, (hi - o) / d
            tn = max(0.0, np.nanmax(np.minimum(t0, t1)))
            tf = min(end - start, np.nanmin(np.maximum(t0, t1)))
            if iv[g, 3] == 1.0:
                if tn > tf + 1e-4:
                    assert dep[g, 7] == 0.0, g                       # missed the root: empty
                    empty += 1
                    continue
                if tf - tn < 1e-3:
                    continue                                         # grazing: too close to call
                s0, s1 = o + d * tn, o + d * tf
                r = max(np.abs(np.concatenate([s0, s1])).max(), 1.0) * 1e-6
                assert dep[g, 7] == 1.0
                assert np.allclose(dep[g, :3], np.minimum(s0, s1) - r, atol=1e-4), (g, dep[g], s0, s1)
                assert np.allclose(dep[g, 4:7], np.maximum(s0, s1) + r, atol=1e-4), (g, dep[g], s0, s1)
                checked += 1
            elif iv[g, 3] == 0.0:
                assert dep[g, 7] == 1.0 and np.all(dep[g, :3] <= dep[g, 4:7])   # a hit is bounded
                hits += 1
    assert checked > 50 and empty > 50 and hits > 10, (checked, empty, hits)


def test_merge_of_a_first_update_with_history_is_the_plain_merge():
    from tests.transport_cases import four_cascade_case
    p, iv, _ = four_cascade_case()
    ns = (iv[:, 3] >= 0).astype(np.uint32)
    blended, hist = O.gi_history(p, 0.88, 0.98, 24, 1, False, iv, ns, np.zeros((iv.shape[0], 16), np.float32),
                                 known_empty(iv.shape[0]))
    a = O.gi_merge(p, iv)
    b = O.gi_merge(p, blended, hist)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_merge_uses_transmittance_without_radiance_mass():
    """A level-0 interval with no radiance mass but transmittance history
    (beta-only blend, no sample this update) passes its children's radiance
    with support beta * mean child support (beginMergeGroup: no own
    radiance; finishMergeGroup: + transmittance * child sums / 4)."""
    from tests.transport_cases import four_cascade_case
    p, iv, idx = four_cascade_case()
    ns = (iv[:, 3] >= 0).astype(np.uint32)
    _, hist = O.gi_history(p, 0.0, 0.9, 24, 1, False, iv, ns, np.zeros((iv.shape[0], 16), np.float32),
                           known_empty(iv.shape[0]))
    g0 = idx["g0"]
    ns2 = ns.copy()
    ns2[g0] = 0
    blended, hist2 = O.gi_history(p, 0.0, 0.9, 24, 1, True, iv, ns2, hist, known_empty(iv.shape[0]))
    assert hist2[g0, 3] == 0.0 and hist2[g0, 5] > 0.0
    m = O.gi_merge(p, blended, hist2)
    plain = O.gi_merge(p, iv)
    tr = blended[g0, 3]
    assert abs(tr - 0.25) < 1e-6
    # plain: J + tr * child (support 0.75 + tr * s_child); here: tr * child only
    child_support = (plain[g0, 3] - 0.75) / 0.25
    assert abs(m[g0, 3] - tr * child_support) < 1e-6
    child_value = (plain[g0, 0] * plain[g0, 3] - 8.0) / (0.25 * child_support)
    assert abs(m[g0, 0] - child_value) < 1e-4
