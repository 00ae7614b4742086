"""Oracle of the temporal interval history (mergeEntry's history part).

Pinned on the reference's own temporal cases (temporalCases,
src/split_merge_tests.rs:31-76, expectations :1303-1359): 512 observations
saturate both masses at 1 / (1 - blend); one opposite observation moves J to
2 j and beta to 1 - b; 200 frames without observations keep value and mass
bit for bit; four observations at once retain j^4 and b^4; each channel's
history is used only when its blend is positive.  Beyond those: the masses
follow the geometric series, J is the geometric-weighted mean of the
observations, a dirty box drops exactly the intervals whose dependency box
meets it, and the merge of a first update with history is the plain merge.
"""
import numpy as np
import pytest

from oracle import oracle as O


def one_interval():
    """A grid with a single interval (one probe, one direction, one level)."""
    return O.gi_params((0.0, 0.0, 0.0), 1.0, (1, 1, 1), 1, base_phi=1, base_y=1)


class Hist:
    """History of the one interval, driven like the reference's advance()."""

    def __init__(self, rb, tb, frames=30, persistent=1):
        self.p = one_interval()
        self.rb, self.tb, self.frames, self.persistent = rb, tb, frames, persistent
        self.h = np.zeros((1, 8), np.float32)
        self.valid = False

    def advance(self, radiance, beta, count):
        iv = np.array([[radiance] * 3 + [beta if count else -1.0]], np.float32)
        out, self.h = O.gi_history(self.p, self.rb, self.tb, self.frames, self.persistent, self.valid, iv,
                                   np.array([count], np.uint32), self.h)
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


def test_dirty_box_drops_exactly_the_intervals_it_meets():
    """Level-0 probe centres x = i + 0.5 (8 x 1 x 1 probes, spacing 1, one
    direction) with interval end 0.4 have dependency boxes [i + 0.1, i +
    0.9] in x and [0.1, 0.9] in y, z; a dirty box at x in [6.0, 6.15] meets
    probe 6's only.  Level 1 is terminal: always dropped.  With emitter bounds
    [0, 0.2]^3 every dependency box reaches back to x = 0, so probes 6 and 7
    (boxes reaching x >= 6.0) are dropped."""
    p = O.gi_params((0.0, 0.0, 0.0), 1.0, (8, 1, 1), 2, base_phi=1, base_y=1, interval_length=0.4,
                    terminal_tail=1)
    n = O.total_intervals(p)
    start, end = O.interval_bounds(p, 0)
    assert abs(end - 0.4) < 1e-6
    iv = np.zeros((n, 4), np.float32)
    iv[:, 0], iv[:, 3] = 3.0, 0.5
    ns = np.ones(n, np.uint32)
    h0 = np.zeros((n, 8), np.float32)
    _, h1 = O.gi_history(p, 0.9, 0.9, 30, 1, False, iv, ns, h0)
    iv2 = iv.copy()
    iv2[:, 0] = 1.0
    box = [((6.0, 0.4, 0.4), (6.15, 0.6, 0.6))]
    out, _ = O.gi_history(p, 0.9, 0.9, 30, 1, True, iv2, ns, h1, boxes=box)
    fresh = out[:, 0] == 1.0
    assert fresh.tolist() == [False] * 6 + [True] + [False] + [True] * (n - 8)
    assert abs(out[0, 0] - (1.0 + 3.0 * 0.9) / 1.9) < 1e-6
    out2, _ = O.gi_history(p, 0.9, 0.9, 30, 1, True, iv2, ns, h1, emit=((0.0, 0.0, 0.0), (0.2, 0.2, 0.2)),
                           boxes=box)
    assert (out2[:8, 0] == 1.0).tolist() == [False] * 6 + [True, True]


def test_merge_of_a_first_update_with_history_is_the_plain_merge():
    from tests.transport_cases import four_cascade_case
    p, iv, _ = four_cascade_case()
    ns = (iv[:, 3] >= 0).astype(np.uint32)
    blended, hist = O.gi_history(p, 0.88, 0.98, 24, 1, False, iv, ns, np.zeros((iv.shape[0], 8), np.float32))
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
    _, hist = O.gi_history(p, 0.0, 0.9, 24, 1, False, iv, ns, np.zeros((iv.shape[0], 8), np.float32))
    g0 = idx["g0"]
    ns2 = ns.copy()
    ns2[g0] = 0
    blended, hist2 = O.gi_history(p, 0.0, 0.9, 24, 1, True, iv, ns2, hist)
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
