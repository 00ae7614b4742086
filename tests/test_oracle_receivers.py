"""Oracle of receiver-driven probe selection and the receiver trace.

Pinned on the reference's own tests:
  * the candidate schedule (src/render.rs:13004-13316): the rank is a
    bijection, each update selects exactly its budget, over a cycle every
    receiver is selected equally often (within one), the 160 x 90 / 512 tail
    cursor values (14336 at batch 28, then (448, 512)), the wrapping frame
    counter, spatial uniformity of low-end batches, the u32 grid limit --
    through the host schedule of the C ABI (lux_candidate_schedule_*) and the
    oracle's rank;
  * the inverse permutation round trip (src/transport_tests.rs:2521-2541);
  * the probe-prefix fixture (src/transport_tests.rs:1887-2480): 64
    candidates, 47 in one cell, distinct keys in one open-addressing
    cluster, negative cells sharing the arithmetic-shift parent, a visible
    receiver without a ray; ray indices are a permutation of the selected
    count, children's ray ranges are contiguous under their parent, roots
    tile [0, selected);
  * receiver-grid and update-budget values of the low-end preset
    (src/render.rs:13316-13336);
  * the cold-priority fixture (src/transport_tests.rs:3015-3620): each
    probe's nominee is its lowest frame-seeded rank, one nominee admitted
    per probe, the fair fill reaching the budget with fair-selected
    receivers only, one LOD per receiver, ordinals 0..count-1 per probe,
    populations 4 and 2 admitted before 1, a budget below the probe count
    admitting exactly that many nominees, warm frames keeping the plain fair
    window; the cold rank's seed through frame 0 and frame = width.
Beyond those: seed keys against floor(p / spacing) per level; the direction
bin against numpy's arctan2 and round trips of the dense bin centres; the
receiver trace's deposits against an independent numpy restatement of the
ray directions and keys, and a closed form (an emissive ceiling: J = the
emission at the level whose interval holds the hit, beta = 1 before it).
"""
import numpy as np
import pytest

from bevy_penumbra_amd import lux
from oracle import oracle as O


def rank_all(w, h):
    n = w * h
    return np.array([O.candidate_rank(s, w, h, n) for s in range(n)])


def selected_set(w, h, budget, cursor, quota):
    rp = O.receiver_params(w, h, budget, cursor, quota)
    return [s for s in range(w * h) if O.candidate_selected(s, rp)]


@pytest.mark.parametrize("w,h,budgets", [(13, 7, [0, 1, 6, 11, 91]), (17, 9, [1, 13, 17, 153]),
                                         (7, 13, [0, 1, 6, 11, 91]), (9, 17, [1, 13, 17, 153])])
def test_schedule_is_bijective_and_exactly_budgeted(w, h, budgets):
    n = w * h
    r = rank_all(w, h)
    assert sorted(r.tolist()) == list(range(n))
    for budget in budgets:
        sched = lux.CandidateSchedule()
        per = n if budget == 0 or budget >= n else budget
        counts = np.zeros(n, np.int64)
        for _ in range(-(-n // per)):
            cursor, quota = sched.phase_for(w, h, n, budget)
            assert quota == min(budget, n)
            sel = selected_set(w, h, budget, cursor, quota)
            assert len(sel) == per
            counts[sel] += 1
            sched.advance_after_successful_update(True)
        assert counts.max() - counts.min() <= 1


def test_schedule_wraps_full_budget_across_uneven_tail():
    W, H, B = 160, 90, 512
    n = W * H
    r = rank_all(W, H)
    sched = lux.CandidateSchedule()
    counts = np.zeros(n, np.int64)
    for batch in range(-(-n // B)):
        cursor, quota = sched.phase_for(W, H, n, B)
        assert quota == B
        if batch == 28:
            assert cursor == 14336
        dist = np.where(r >= cursor, r - cursor, n - cursor + r)
        sel = dist < quota
        # the oracle's per-sample selection agrees with the vectorised window
        for s in (0, 1, 777, n - 1):
            assert O.candidate_selected(s, O.receiver_params(W, H, B, cursor, quota)) == bool(sel[s])
        assert sel.sum() == B
        counts[sel] += 1
        sched.advance_after_successful_update(True)
    assert (counts.min(), counts.max()) == (1, 2)
    assert sched.phase_for(W, H, n, B) == (448, B)


def test_low_end_batches_are_spatially_uniform():
    W, H, B, T = 160, 90, 1024, 8
    n = W * H
    r = rank_all(W, H)
    sched = lux.CandidateSchedule()
    covered = np.zeros(n, bool)
    while not covered.all():
        cursor, quota = sched.phase_for(W, H, n, B)
        dist = np.where(r >= cursor, r - cursor, n - cursor + r)
        sel = np.nonzero(dist < quota)[0]
        covered[sel] = True
        x, y = sel % W, sel // W
        tiles = np.zeros((-(-H // T), -(-W // T)), np.int64)
        np.add.at(tiles, (y // T, x // T), 1)
        assert len(sel) <= B
        if quota == B:
            assert len(sel) == B
            assert (tiles > 0).sum() >= 225
            assert (tiles.sum(axis=0) > 0).all() and (tiles.sum(axis=1) > 0).all()
            assert tiles.max() <= 12
        sched.advance_after_successful_update(True)


def test_schedule_phase_survives_wrapping_frame_counter_and_u32_grid():
    sched = lux.CandidateSchedule()
    for expected in [0, 3, 6, 9, 12]:
        c, q = sched.phase_for(5, 3, 15, 3)
        assert (c, q) == (expected, 3)
        assert sched.phase_for(5, 3, 15, 3) == (c, q)      # preparing does not consume a phase
        sched.advance_after_successful_update(True)
    assert sched.phase_for(5, 3, 15, 3)[0] == 0
    w, h = 65535, 65537
    n = w * h
    assert n == 0xFFFFFFFF
    for s in [0, 1, w - 1, w, n - 2, n - 1]:
        assert O.candidate_rank(s, w, h, n) < n
    big = lux.CandidateSchedule()
    assert big.phase_for(w, h, n, 0x80000000) == (0, 0x80000000)
    big.advance_after_successful_update(True)
    assert big.phase_for(w, h, n, 0x80000000) == (0x80000000, 0x80000000)
    big.advance_after_successful_update(True)
    assert big.phase_for(w, h, n, 0x80000000) == (1, 0x80000000)
    # a submitted update that does not consume the window keeps the cursor
    s2 = lux.CandidateSchedule()
    s2.phase_for(160, 90, 14400, 512)
    s2.advance_after_successful_update(False)
    assert s2.phase_for(160, 90, 14400, 512) == (0, 512)


def test_inverse_permutation_round_trips_irregular_counts():
    for count in [1, 3, 5, 7, 9, 17, 31, 33, 63, 65, 127, 129]:
        seen = set()
        for rank in range(count):
            s = O.candidate_from_rank(rank, count, 1, count)
            assert s not in seen and s < count
            seen.add(s)
            assert O.candidate_rank(s, count, 1, count) == rank
        assert len(seen) == count


def test_receiver_grid_and_budget_of_the_low_end_preset():
    le = lux.RadianceCascades.low_end()
    assert le.receiver_grid(1280, 720) == (160, 90)
    assert le.update_budget() == 1024
    assert le.update_budget(32) == 32
    assert le.update_budget(2048) == 1024
    assert le.update_budget(0) == 1
    one = lux.RadianceCascades.low_end()
    one.max_rays = 1
    assert one.receiver_grid(1280, 720) == (160, 90) and one.update_budget() == 1
    capped = lux.RadianceCascades.low_end()
    capped.sampling[1] = 256
    w, h = capped.receiver_grid(1280, 720)
    assert w * h <= 256 and capped.update_budget() == 1024


def test_disabled_and_zero_budget_selection():
    rp = O.receiver_params(4, 4, 0)
    assert all(O.candidate_selected(s, rp) for s in range(16))       # budget 0: every receiver
    rp = O.receiver_params(4, 4, 20)
    assert all(O.candidate_selected(s, rp) for s in range(16))       # budget >= count
    rp = O.receiver_params(4, 4, 0, disabled=True)
    assert not any(O.candidate_selected(s, rp) for s in range(16))
    rp = O.receiver_params(4, 4, 5, fair_quota=0)
    assert not any(O.candidate_selected(s, rp) for s in range(16))


# ---------------- the probe-prefix fixture ----------------
def prefix_fixture():
    cands, levels = 64, 4
    cap = 1 << (cands * 4 - 1).bit_length()

    def probe_bucket(cell, lod, level):
        m = 0xFFFFFFFF
        h = ((cell & m) * 0x8DA6B343 & m) ^ ((((lod << 18) | (level << 12)) * 0x165667B1) & m)
        h ^= h >> 16
        h = (h * 0x7FEB352D) & m
        h ^= h >> 15
        return h & (cap - 1)

    target = probe_bucket(0, 0, 0)
    coll, roots = [], []
    for cell in range(1, 1_000_000):
        if probe_bucket(cell, 0, 0) == target and (cell >> 3) not in roots:
            coll.append(cell)
            roots.append(cell >> 3)
            if len(coll) == 4:
                break
    assert len(coll) == 4
    unique = list(coll)
    for cell in [8, 16, 24, 32, 40, 48, 56, 64, 72, 80, 88, 96]:
        if cell not in unique:
            unique.append(cell)
    unique = unique[:15]
    cells = [0] * cands
    cells[0] = 1_000_000
    cells[1:48] = [1_000_003] * 47
    cells[48:63] = unique[:15]
    cells[61] = -2
    cells[62] = -1
    cells[63] = 1_000_007
    nd = cands * 2 * levels
    keys = np.zeros((nd, 4), np.int32)
    flags = np.zeros(nd, np.uint32)
    for s, cell in enumerate(cells):
        for v in range(2):
            for L in range(levels):
                d = (v * cands + s) * levels + L
                keys[d] = (cell >> L, 0, 0, (v << 18) | (L << 12))
                flags[d] = O.RX_VALID | (O.RX_RAY if (L == 0 and s < 48 and v == 0) else 0)
    return cands, levels, cells, keys, flags, cap


def test_prefix_fixture_matches_the_reference_expectations():
    cands, levels, cells, keys, flags, cap = prefix_fixture()
    assert O._tsig().orc_rx_directory_capacity(cands) == cap == 256
    r = O.rx_select(cands, levels, keys, flags)
    nd = keys.shape[0]
    assert (r["canonical"] != O.EMPTY).all() and (r["canonical"] < nd).all()
    selected = 48
    seq = r["sequence"]
    rays = sorted(seq[0][seq[0] != O.EMPTY].tolist())
    assert rays == list(range(selected))
    assert (seq[1] == O.EMPTY).all()                               # the overlap traced only its low LOD
    assert r["root"] == selected
    assert r["tickets"].tolist() == list(range(selected))
    assert r["canonical"][0] == 0 and r["offset"][0] == 0          # first child: valid zero prefix
    nodes = {}
    for s, cell in enumerate(cells):
        for v in range(2):
            for L in range(levels):
                d = (v * cands + s) * levels + L
                node = (v, L, cell >> L)
                got = (int(r["canonical"][d]), int(r["offset"][d]), int(r["count"][d]))
                assert nodes.setdefault(node, got) == got, node     # one probe per grid key
    assert nodes[(0, 0, 1_000_007)][2] == 0                        # visible, never selected
    assert nodes[(0, 1, -1)][2] == 0
    assert sorted(c for (v, L, c) in nodes if v == 0 and L == 0 and c >> 1 == -1) == [-2, -1]
    assert nodes[(0, 0, 1_000_003)][2] == 47
    for (v, L, c), (_, off, cnt) in nodes.items():
        if L == 0:
            continue
        kids = sorted((cc, nodes[(vv, LL, cc)]) for (vv, LL, cc) in nodes if vv == v and LL == L - 1 and cc >> 1 == c)
        cursor = off
        for cc, (_, koff, kcnt) in kids:
            assert koff == cursor, (v, L, c, cc)
            cursor += kcnt
        assert cursor == off + cnt
    roots = sorted((off, cnt) for (v, L, c), (_, off, cnt) in nodes.items() if L == 3 and cnt > 0)
    cursor = 0
    for off, cnt in roots:
        assert off == cursor
        cursor += cnt
    assert cursor == selected and roots[0][0] == 0


def test_prefix_fixture_orders_are_the_sequential_ones():
    """Canonical = lowest descriptor index of its key; ordinals ascend with
    the receiver index inside a probe (47 receivers of one cell get
    consecutive rays in receiver order)."""
    cands, levels, cells, keys, flags, _ = prefix_fixture()
    r = O.rx_select(cands, levels, keys, flags)
    canon = r["canonical"]
    for d in range(keys.shape[0]):
        same = np.nonzero((keys[:, :3] == keys[d, :3]).all(axis=1) & (keys[:, 3] == keys[d, 3]))[0]
        assert canon[d] == same.min()
    s47 = r["sequence"][0][1:48]
    assert (np.diff(s47.astype(np.int64)) == 1).all()


# ---------------- cold priority ----------------
def test_cold_rank_folds_the_frame_into_the_permutation_seed():
    """coldCandidateRank (split.slang:631-657, restated by the reference at
    src/transport_tests.rs:1170-1194): the fair permutation with seed
    (w*phi ^ h*c ^ count ^ frame*phi) & mask.  Frame 0 gives the fair rank;
    frame = width cancels the width term, which is the fair rank of width 0;
    every frame is a bijection with the inverse as its inverse."""
    for w, h in [(8, 1), (13, 7), (160, 90), (31, 3)]:
        n = w * h
        for frame in [0, 1, 128, 977, w]:
            ranks = [O.cold_candidate_rank(s, w, h, n, frame) for s in range(n)]
            assert sorted(ranks) == list(range(n))
            for s in range(0, n, max(1, n // 50)):
                assert O.cold_candidate_from_rank(ranks[s], w, h, n, frame) == s
        fair = [O.candidate_rank(s, w, h, n) for s in range(n)]
        assert [O.cold_candidate_rank(s, w, h, n, 0) for s in range(n)] == fair
        assert [O.cold_candidate_rank(s, w, h, n, w) for s in range(n)] == \
            [O.candidate_rank(s, 0, h, n) for s in range(n)]
    # frame * phi reaches the 3-bit seed of 8 receivers only through odd frames
    fair8 = [O.candidate_rank(s, 8, 1, 8) for s in range(8)]
    assert [O.cold_candidate_rank(s, 8, 1, 8, 128) for s in range(8)] == fair8
    assert [O.cold_candidate_rank(s, 8, 1, 8, 1) for s in range(8)] != fair8


COLD_W, COLD_FRAME = 8, 128


def cold_fixture(weighted):
    """The entries of src/transport_tests.rs:3015-3620: 8 receivers on one
    row, 4 levels, keys (group >> level, 0, 0, variant << 18 | level << 12),
    the selected LOD = receiver % 2.  Unweighted: group = receiver / 4 (four
    probes of population 2).  Weighted: variant 0 is group 0; variant 1 is
    group 0 for receivers 1 and 3, 1 for 5, 2 for 7, 3 otherwise."""
    count, levels = COLD_W, 4
    keys = np.zeros((2 * count * levels, 4), np.int32)
    flags = np.zeros(2 * count * levels, np.uint32)
    for s in range(count):
        for v in range(2):
            if not weighted:
                group = s // 4
            else:
                group = 0 if v == 0 else {1: 0, 3: 0, 5: 1, 7: 2}.get(s, 3)
            for L in range(levels):
                d = (v * count + s) * levels + L
                keys[d] = (group >> L, 0, 0, (v << 18) | (L << 12))
                flags[d] = O.RX_VALID | (O.RX_RAY if L == 0 and v == s % 2 else 0)
    return count, levels, keys, flags


def cold_select(weighted, budget):
    count, levels, keys, flags = cold_fixture(weighted)
    rp = O.receiver_params(COLD_W, 1, budget, 2, budget, frame=COLD_FRAME)
    return count, levels, keys, rp, O.rx_select(count, levels, keys, flags, rp, cold=True)


def traced(r, count, levels):
    """(receiver, variant) -> local ordinal of every traced ray."""
    out = {}
    for v in range(2):
        for s in range(count):
            q = int(r["sequence"][v, s])
            if q != O.EMPTY:
                out[(s, v)] = q - int(r["offset"][(v * count + s) * levels])
    return out


def probe_of(r, count, levels, s, v):
    return int(r["canonical"][(v * count + s) * levels])


def test_cold_priority_nominates_each_probe_then_fills_the_fair_budget():
    count, levels, keys, rp, r = cold_select(False, 6)
    fair = [O.candidate_selected(s, rp) for s in range(count)]
    rays = traced(r, count, levels)
    assert len(rays) == 6 and r["root"] == 6                       # fair fill reaches the budget
    assert len({s for s, _ in rays}) == 6                          # one LOD per receiver
    per_probe = {}
    for (s, v), ordinal in rays.items():
        per_probe.setdefault(probe_of(r, count, levels, s, v), []).append((s, ordinal))
    for g in range(2):
        for v in range(2):
            first, second = g * 4 + v, g * 4 + v + 2
            nominee = first if O.cold_candidate_rank(first, COLD_W, 1, count, COLD_FRAME) < \
                O.cold_candidate_rank(second, COLD_W, 1, count, COLD_FRAME) else second
            probe = probe_of(r, count, levels, first, v)
            assert probe == (v * count + g * 4) * levels           # canonical: lowest descriptor
            assert rays[(nominee, v)] == 0                         # every probe's nominee, ordinal 0
            got = per_probe[probe]
            assert sorted(o for _, o in got) == list(range(len(got))) == list(range(int(r["count"][probe])))
            for s, o in got:
                if s != nominee:
                    assert fair[s] and o > 0                       # fill receivers are fair-selected
    assert ((r["flags"][::levels] & 4) != 0).sum() == 4                # one admitted nominee per probe
    # the fill takes the fair-selected non-nominees in ascending candidate rank
    nominees = {s for (s, v), o in rays.items() if o == 0}
    fill = sorted((O.candidate_rank(s, COLD_W, 1, count), s) for s in range(count) if fair[s] and s not in nominees)
    assert {s for _, s in fill[:2]} == {s for (s, v) in rays if s not in nominees}


def test_cold_budget_admits_dense_probes_before_sparse_ones():
    count, levels, keys, rp, r = cold_select(True, 2)
    # projected populations of the selected-LOD receivers' probes
    pop = {}
    for s in range(count):
        pop[tuple(keys[((s % 2) * count + s) * levels])] = pop.get(tuple(keys[((s % 2) * count + s) * levels]), 0) + 1
    assert [pop[tuple(keys[((s % 2) * count + s) * levels])] for s in range(count)] == [4, 2, 4, 2, 4, 1, 4, 1]
    counts = [int(r["count"][((s % 2) * count + s) * levels]) for s in [0, 1, 5, 7]]
    assert counts == [1, 1, 0, 0]                                  # populations 4 and 2 before 1
    assert r["root"] == 2 and len(traced(r, count, levels)) == 2


def test_cold_admission_stays_within_a_budget_below_the_probe_count():
    count, levels, keys, rp, r = cold_select(False, 2)
    rays = traced(r, count, levels)
    assert len(rays) == 2 and r["root"] == 2
    assert sorted(int(r["count"][(v * count + g * 4 + v) * levels]) for g in range(2) for v in range(2)) == [0, 0, 1, 1]
    assert all(o == 0 for o in rays.values())                      # each ray is its probe's nominee
    assert ((r["flags"][::levels] & 4) != 0).sum() == 2
    # admission tickets inside one population bin go in the rank-mapped
    # dispatch order of the canonical descriptors
    order = sorted((O.candidate_rank(g * 4, COLD_W, 1, count) * 2 + v, (g, v)) for g in range(2) for v in range(2))
    admitted = {(s // 4, v) for (s, v) in rays}
    assert admitted == {gv for _, gv in order[:2]}


def test_warm_selection_is_the_plain_fair_window():
    count, levels, keys, flags = cold_fixture(False)
    rp = O.receiver_params(COLD_W, 1, 6, 2, 6, frame=COLD_FRAME)
    fair = [O.candidate_selected(s, rp) for s in range(count)]
    warm_flags = flags.copy()
    for s in range(count):
        if not fair[s]:
            warm_flags[((s % 2) * count + s) * levels] &= ~np.uint32(O.RX_RAY)
    r = O.rx_select(count, levels, keys, warm_flags)
    rays = traced(r, count, levels)
    assert sorted(rays) == sorted((s, s % 2) for s in range(count) if fair[s])
    assert len(rays) == 6 and r["root"] == 6


def test_cold_seed_makes_every_receiver_eligible():
    rng = np.random.default_rng(5)
    w, h, levels = 6, 5, 3
    n = w * h
    rc = np.zeros((n, 6), np.float32)
    rc[:, :3] = rng.uniform(-2, 2, (n, 3))
    rc[:, 4] = 1.0
    rc[7, 3:] = 0.0                                                # no surface
    rp = O.receiver_params(w, h, 4, 3, 4, frame=9)
    keys, flags, _ = O.rx_seed(rc, rp, None, 0.5, levels)
    ckeys, cflags, _ = O.rx_seed(rc, rp, None, 0.5, levels, cold=True)
    assert (keys == ckeys).all()
    ray = (flags[: n * levels: levels] & O.RX_RAY) != 0
    cray = (cflags[: n * levels: levels] & O.RX_RAY) != 0
    assert ray.sum() == 4 and cray.sum() == n - 1 and not cray[7] and (cray | ~ray).all()


# ---------------- seed ----------------
def test_seed_keys_lods_and_flags():
    rng = np.random.default_rng(3)
    w, h, levels, spacing = 8, 6, 4, 0.25
    n = w * h
    rc = np.zeros((n, 6), np.float32)
    rc[:, :3] = rng.uniform(-3, 3, (n, 3))
    rc[:, 4] = 1.0
    rc[5, 3:] = 0.0                                     # no surface
    rc[6, 0] = np.inf                                   # non-finite position
    view = O.lod_view(camera=(0, 0, 0), probe_spacing=spacing, clip_row1=(0, 1, 0), viewport_height=48,
                      clip_row3=(0, 0, -1), target_pixels=8)
    rp = O.receiver_params(w, h, 0, frame=5)
    keys, flags, lods = O.rx_seed(rc, rp, view, spacing, levels)
    keys = keys.reshape(2, n, levels, 4)
    flags = flags.reshape(2, n, levels)
    saw_overlap = False
    for s in range(n):
        if s in (5, 6):
            assert (flags[:, s] == 0).all()
            continue
        lo, hi, blend = O.lod_select(view, rc[s, :3])
        assert (lods[0, s], lods[1, s]) == (lo, hi)
        for v, lod in ((0, lo), (1, hi)):
            if v == 1 and hi == lo:
                assert (flags[1, s] == 0).all()
                continue
            for L in range(levels):
                sp = np.float32(spacing) * np.float32(2.0 ** (L + lod))
                want = np.floor(rc[s, :3] / sp).astype(np.int32)
                assert (keys[v, s, L, :3] == want).all()
                assert keys[v, s, L, 3] == (L << 12) | (lod << 18)
        rays = flags[:, s, 0] & O.RX_RAY
        assert rays.sum() == O.RX_RAY                    # budget 0: exactly one variant traces
        assert not (flags[:, s, 1:] & O.RX_RAY).any()
        if hi != lo:
            saw_overlap = True
            bits = (s ^ (5 * 0x9E3779B9 & 0xFFFFFFFF)) & 0xFFFFFFFF
            bits ^= bits >> 16
            bits = bits * 0x7FEB352D & 0xFFFFFFFF
            bits ^= bits >> 15
            bits = bits * 0x846CA68B & 0xFFFFFFFF
            bits ^= bits >> 16
            phase = np.float32(bits >> 8) * np.float32(1.0 / 16777216.0)
            assert (flags[1, s, 0] & O.RX_RAY != 0) == bool(phase < np.float32(blend))
    assert saw_overlap
    _, fl, _ = O.rx_seed(rc, O.receiver_params(w, h, 0, disabled=True), view, spacing, levels)
    assert not (fl & O.RX_RAY).any()


# ---------------- direction bins ----------------
def test_direction_index_against_arctan2_and_bin_centres():
    rng = np.random.default_rng(1)
    d = rng.normal(size=(4000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    for L in range(4):
        nphi, ny = 8 << L, 4 << L
        u = np.mod(np.arctan2(d[:, 2].astype(np.float64), d[:, 0].astype(np.float64)) / (2 * np.pi) + 1.0, 1.0)
        v = np.clip(d[:, 1].astype(np.float64) * 0.5 + 0.5, 0, 1)
        x = np.minimum((u * nphi).astype(int), nphi - 1)
        y = np.minimum((v * ny).astype(int), ny - 1)
        want = x + y * nphi
        got = np.array([O.direction_index(di, nphi, ny) for di in d])
        near = (np.abs(u * nphi - np.round(u * nphi)) < 1e-5) | (np.abs(v * ny - np.round(v * ny)) < 1e-5)
        assert (got[~near] == want[~near]).all()
        for k in range(nphi * ny):                       # the dense grid's bin centres come back
            assert O.direction_index(O.bin_direction(k, nphi, ny), nphi, ny) == k
    assert O.direction_index((-1.0, 0.0, -0.0), 8, 4) == 4 + 2 * 8    # atan2(-0, -1) = -pi: u = 0.5
    assert O.direction_index((0.0, -1.0, 0.0), 8, 4) == 0


# ---------------- the receiver trace ----------------
def ceiling_scene(height=1.3):
    """An emissive ceiling (emission (2, 3, 4), albedo 0) over a receiver
    plane; nothing else."""
    from bevy_penumbra_amd import scenes
    quad = scenes.quad((-500.0, height, -500.0), (0, 0, 1000), (1000, 0, 0)).astype(np.float32)
    sc = scenes.Scene(meshes=[quad], mesh_materials=[np.zeros(quad.shape[0], np.uint32)],
                      inst_inv=np.eye(4, dtype=np.float32)[:3].reshape(1, 12),
                      inst_fwd=np.eye(4, dtype=np.float32)[:3].reshape(1, 12),
                      inst_mesh=np.zeros(1, np.uint32), inst_flags=np.zeros(1, np.uint32),
                      materials=np.array([[0, 0, 0, 1, 2, 3, 4, 0]], np.float32),
                      bounds=((-500, 0, -500), (500, height, 500)))
    return sc


def r2(i):
    i = np.asarray(i, np.uint64)
    px = (i * np.uint64(3242174889)) & np.uint64(0xFFFFFFFF)
    py = (i * np.uint64(2447445414)) & np.uint64(0xFFFFFFFF)
    return ((px >> np.uint64(8)).astype(np.float32) / np.float32(16777216.0),
            (py >> np.uint64(8)).astype(np.float32) / np.float32(16777216.0))


def thash(v):
    v = int(v) & 0xFFFFFFFF
    v ^= v >> 16
    v = v * 0x7FEB352D & 0xFFFFFFFF
    v ^= v >> 15
    v = v * 0x846CA68B & 0xFFFFFFFF
    v ^= v >> 16
    return v


def test_receiver_trace_deposits_match_a_numpy_restatement_and_the_closed_form():
    height = 1.3
    sc = ceiling_scene(height)
    built = O.build_scene(sc)
    lights = O.build_lights(sc)
    w, h = 12, 10
    n = w * h
    rng = np.random.default_rng(5)
    rc = np.zeros((n, 6), np.float32)
    rc[:, 0] = rng.uniform(0.05, 1.95, n)
    rc[:, 1] = 0.02
    rc[:, 2] = rng.uniform(0.05, 1.95, n)
    rc[:, 4] = 1.0
    levels, spacing, frame = 4, 0.25, 9
    g = O.gi_params((0.0, 0.0, 0.0), spacing, (8, 8, 8), levels, base_phi=8, base_y=4, interval_length=0.4,
                    emissive_samples=4, frame=frame, terminal_tail=1)
    rp = O.receiver_params(w, h, 0, frame=frame)
    keys, flags, lods = O.rx_seed(rc, rp, None, spacing, levels)
    sel = O.rx_select(n, levels, keys, flags)
    ivs, sms, st = O.rx_trace(built, lights, sc.materials, [g], rc, lods, sel["sequence"], sel["tickets"], frame)
    iv, sm = ivs[0], sms[0]
    assert st[0] == n and st[1] == 0
    # numpy restatement of the ray directions, keys and the deposit walk
    jx, jy = r2(thash(frame ^ 0x5A17C9E3))
    seq = sel["sequence"][0].astype(np.uint64)
    qx, qy = r2(seq + 1)
    ux = (np.float32(0.5) + qx) + jx
    uy = (np.float32(0.5) + qy) + jy
    ux, uy = ux - np.floor(ux), uy - np.floor(uy)
    cs = np.array([O.cos_sin_2pi(u) for u in ux], np.float32)
    y = uy * np.float32(2) - np.float32(1)
    rr = np.sqrt(np.maximum(np.float32(0), np.float32(1) - y * y))
    d = np.stack([rr * cs[:, 0], y, rr * cs[:, 1]], 1).astype(np.float32)
    d[d[:, 1] < 0] *= -1                                # flipped into the receiver's hemisphere
    oy = (np.float32(0.02) + np.float32(0.005)) + np.float32(1.0 / 65536.0)   # rayOffset below 1/32
    t = (np.float32(height) - oy) / d[:, 1]
    hit = np.all(np.abs(rc[:, [0, 2]] + d[:, [0, 2]] * t[:, None]) < 500.0, axis=1)
    expect = {}
    for i in range(n):
        for L in range(levels):
            start = 0.0 if L == 0 else 0.4 * 4.0 ** (L - 1)
            end = 0.4 * 4.0 ** L
            term = L == levels - 1 or (hit[i] and start < t[i] <= end)
            cell = np.floor(rc[i, :3].astype(np.float64) / (spacing * 2 ** L)).astype(int)
            dims = 8 >> L
            nphi, ny = 8 << L, 4 << L
            u = np.mod(np.arctan2(d[i, 2], d[i, 0]) / (2 * np.pi) + 1.0, 1.0)
            k = min(int(u * nphi), nphi - 1) + min(int((d[i, 1] * 0.5 + 0.5) * ny), ny - 1) * nphi
            probe = cell[0] + dims * (cell[1] + dims * cell[2])
            slot = O.level_offset(g, L) + probe * nphi * ny + k
            expect.setdefault(slot, []).append((L, term and hit[i]))
            if term:
                break
    got = np.nonzero(sm)[0]
    assert sorted(expect) == got.tolist()
    for slot, deps in expect.items():
        assert sm[slot] == len(deps)
        n_hit = sum(1 for L, tm in deps if tm)
        n_pass = sum(1 for L, tm in deps if L != levels - 1 and not tm)
        # closed form: a ray ending on the ceiling carries its emission (albedo 0: no NEE term), a
        # ray passing the interval carries beta 1, an escape in the terminal tail J = 0, beta 0
        assert np.allclose(iv[slot, :3], np.array([2, 3, 4]) * n_hit / len(deps), atol=1e-6)
        assert abs(iv[slot, 3] - n_pass / len(deps)) < 1e-6
    assert (iv[sm == 0, 3] == -1).all() and (iv[sm == 0, :3] == 0).all()
    assert st[2] == sum(len(v) for v in expect.values()) and st[3] == 0


def test_receiver_trace_drops_deposits_outside_the_grid_and_lods_without_grid():
    sc = ceiling_scene()
    built = O.build_scene(sc)
    lights = O.build_lights(sc)
    n = 16
    rc = np.zeros((n, 6), np.float32)
    rc[:, 0] = np.linspace(-1.0, 3.0, n)                 # half of them outside x in [0, 2)
    rc[:, 1] = 0.02
    rc[:, 2] = 0.5
    rc[:, 4] = 1.0
    g = O.gi_params((0.0, 0.0, 0.0), 0.25, (8, 8, 8), 2, interval_length=0.4, emissive_samples=4)
    keys, flags, lods = O.rx_seed(rc, O.receiver_params(4, 4, 0), None, 0.25, 2)
    sel = O.rx_select(n, 2, keys, flags)
    lods2 = lods.copy()
    lods2[0, :4] = 3                                     # a LOD the pyramid does not hold
    ivs, sms, st = O.rx_trace(built, lights, sc.materials, [g], rc, lods2, sel["sequence"], sel["tickets"], 0)
    inside = (rc[:, 0] >= 0) & (rc[:, 0] < 2.0)
    inside[:4] = False
    assert st[0] == n
    assert sms[0][:O.level_offset(g, 1)].sum() == inside.sum()    # one level-0 deposit per kept ray
    assert st[3] > 0 and st[2] + st[3] >= n
