"""Batched GI update of uneven scenes (lux_batch) against each scene's own update.

Five clutter rooms of different triangle counts, probe grids, cascade counts,
rays per interval and NEE sample counts run as one batch: every scene's
intervals, merged radiance, packed probes and query counters must equal, bit
for bit, what the scene's own lux_gi_update writes (which the transport tests
hold to the oracle).  A scene with a surrogate must equal its own
lux_gi_update_surrogate with no fallback, and with a finite error bound (its
traced per-probe fallback inside the batch launch, eager and replayed) its
own update with that bound.  The captured graph must give the
eager result on every replay, with eager work between replays, and pick up a
new frame set after capture.
"""
import ctypes

import numpy as np
import pytest

from bevy_penumbra_amd import lux, scenes

pytestmark = pytest.mark.gpu

SPECS = [
    # triangles, room side, grid side, levels, rays per interval, emissive samples
    (300, 4.0, 4, 2, 1, 4),
    (6000, 6.0, 10, 4, 1, 16),
    (1500, 5.0, 7, 3, 2, 4),
    (90, 4.0, 3, 2, 1, 2),
    (3000, 8.0, 8, 3, 1, 8),
]


def make_scene(i, spec):
    tri, size, d, levels, R, ns = spec
    sc, _ = scenes.clutter_room(100 + i, tri, size)
    ctx = lux.Context(0)
    ctx.upload_scene(sc)
    ctx.build_bvh()
    sp = size / d
    ctx.gi_configure(lux.gi_params((0.5 * sp, 0.5 * sp, 0.5 * sp), sp, (d, d, d), levels, base_phi=4, base_y=4,
                                   interval_length=sp, emissive_samples=ns, rays_per_interval=R,
                                   seed=i, terminal_tail=1))
    return ctx


def poison(ctx):
    """Overwrite the scene's interval, merged and probe arrays with NaN bits."""
    n, probes, _ = ctx.gi_sizes()
    for ptr, nbytes in zip(ctx.gi_device_arrays(), (16 * n, 16 * n, 16 * 36 * probes)):
        bad = np.full(nbytes // 4, 0xFFFFFFFF, np.uint32)
        ctx._check(ctx.lib.lux_copy_to_device(ctx.h, ptr, bad.ctypes.data, nbytes), "lux_copy_to_device")
    ctx.synchronize()


def state(ctx):
    out = ctx.gi_download()
    out["counters"] = ctx.gi_query_counters()
    return out


def assert_same(a, b, what, upto=None):
    for key in ("intervals", "merged", "probes"):
        x, y = a[key], b[key]
        if upto is not None and key != "probes":
            x, y = x[:upto[key]], y[:upto[key]]
        assert np.array_equal(np.ascontiguousarray(x).view(np.uint32),
                              np.ascontiguousarray(y).view(np.uint32)), f"{what}: {key}"
    assert a["counters"] == b["counters"], f"{what}: counters {a['counters']} vs {b['counters']}"


@pytest.fixture(scope="module")
def rooms():
    ctxs = [make_scene(i, s) for i, s in enumerate(SPECS)]
    yield ctxs
    for c in ctxs:
        c.close()


def per_scene(ctxs, frame, seed):
    refs = []
    for c in ctxs:
        c.gi_set_frame(frame, seed)
        c.gi_update()
        c.synchronize()
        refs.append(state(c))
    return refs


def batch_states(ctxs, b):
    b.synchronize()
    return [state(c) for c in ctxs]


def test_gpu_batch_equals_per_scene_updates(rooms):
    refs = per_scene(rooms, 7, 3)
    for c in rooms:
        poison(c)
    b = lux.Batch(rooms)
    sz = b.sizes()
    assert sz["scenes"] == len(SPECS)
    assert sz["intervals"] == sum(c.gi_sizes()[0] for c in rooms)
    b.set_frame(7, 3)
    b.update()
    for i, (r, g) in enumerate(zip(refs, batch_states(rooms, b))):
        assert_same(r, g, f"scene {i}")
    # another launch configuration and a frame change give the same bits as
    # the per-scene updates of that frame
    refs8 = per_scene(rooms, 8, 9)
    b.set_config(waves_per_cu=24, lds_stack_depth=4, regroup_threshold=4, steps_per_refill=2)
    for c in rooms:
        poison(c)
    b.set_frame(8, 9)
    b.update()
    for i, (r, g) in enumerate(zip(refs8, batch_states(rooms, b))):
        assert_same(r, g, f"scene {i}, frame 8, config 2")
    b.set_config(waves_per_cu=20, lds_stack_depth=12, regroup_threshold=16, steps_per_refill=16)
    for c in rooms:
        poison(c)
    b.update()
    for i, (r, g) in enumerate(zip(refs8, batch_states(rooms, b))):
        assert_same(r, g, f"scene {i}, frame 8, config 3")
    b.close()


def test_gpu_batch_graph_replays_match_eager(rooms):
    b = lux.Batch(rooms)
    b.set_frame(11, 5)
    b.update()
    eager = batch_states(rooms, b)
    b.capture()
    for rep in range(3):
        for c in rooms:
            poison(c)
        b.launch()
        got = batch_states(rooms, b)
        for i, (r, g) in enumerate(zip(eager, got)):
            assert_same(r, g, f"replay {rep} scene {i}")
        # eager work on the scene contexts between replays
        rooms[1].gi_set_frame(2, 2)
        rooms[1].gi_update()
        rooms[1].synchronize()
    # a frame set after capture is picked up by the graph
    refs = per_scene(rooms, 12, 6)
    for c in rooms:
        poison(c)
    b.set_frame(12, 6)
    b.launch()
    for i, (r, g) in enumerate(zip(refs, batch_states(rooms, b))):
        assert_same(r, g, f"graph at frame 12, scene {i}")
    ms = b.launch(timed=True)
    assert ms > 0
    b.close()


def test_gpu_batch_surrogate_scene_equals_its_own_update(rooms):
    """Scene 1 predicts level 2 with an (untrained) network, scene 4 level 1;
    each equals lux_gi_update_surrogate with error_bound = +inf; the other
    scenes stay exactly their plain update."""
    sur = {}
    for i, level in ((1, 2), (4, 1)):
        c = rooms[i]
        tri, size, d, levels, R, ns = SPECS[i]
        sur[i] = (lux.Surrogate(c, lux.SurrogateParams.default((0, 0, 0), (size, size, size), seed=i + 1)), level)
    refs = per_scene(rooms, 4, 4)
    uptos = {}
    for i, (s, level) in sur.items():
        c = rooms[i]
        st = c.gi_update_surrogate(s, level, float("inf"))
        assert st.fallback_probes == 0
        c.synchronize()
        refs[i] = state(c)
        offs = level_offsets(c)
        uptos[i] = dict(intervals=offs[level], merged=offs[level + 1])
    for c in rooms:
        poison(c)
    b = lux.Batch(rooms, surrogates=sur)
    b.set_frame(4, 4)
    b.update()
    got = batch_states(rooms, b)
    for i in range(len(rooms)):
        assert_same(refs[i], got[i], f"scene {i}", upto=uptos.get(i))
    # the graph reads the network's weights at replay time: a training step
    # after capture changes the prediction exactly as the eager update does
    b.capture()
    s, level = sur[4]
    samples = lux.DeviceBuffer.empty(rooms[4], (4096, 12), np.float32)
    rooms[4].gi_surrogate_samples(level, 4096, 1, lux.LUX_SAMPLES_TRACE, samples)
    s.train_step(samples, 4096, want_loss=False)
    rooms[4].synchronize()
    rooms[4].gi_update_surrogate(s, level, float("inf"), stats=False)
    rooms[4].synchronize()
    ref4 = state(rooms[4])
    poison(rooms[4])
    b.launch()
    got4 = batch_states(rooms, b)[4]
    assert_same(ref4, got4, "scene 4 after a training step", upto=uptos[4])
    b.close()
    for s, _ in sur.values():
        s.close()


def test_gpu_batch_surrogate_fallback_equals_its_own_update(rooms):
    """A finite error bound: scene 4's probes above it are traced inside the
    batch launch (eager and graph replay), equal to lux_gi_update_surrogate
    with the same bound; a bound of 0 traces every probe of the level."""
    i, level = 4, 1
    c = rooms[i]
    tri, size, d, levels, R, ns = SPECS[i]
    s = lux.Surrogate(c, lux.SurrogateParams.default((0, 0, 0), (size, size, size), seed=7))
    refs = per_scene(rooms, 4, 4)
    # a few training steps make the predicted errors differ between probes
    samples = lux.DeviceBuffer.empty(c, (4096, 12), np.float32)
    for k in range(3):
        c.gi_surrogate_samples(level, 4096, k + 1, lux.LUX_SAMPLES_TRACE, samples)
        s.train_step(samples, 4096, want_loss=False)
    c.synchronize()
    st = c.gi_update_surrogate(s, level, 0.0)
    assert st.fallback_probes == st.probes                          # bound 0: every probe traced
    lo, hi = 0.0, 1e6                                               # bisect to a partial fallback
    chosen = None
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        nf = c.gi_update_surrogate(s, level, mid).fallback_probes
        if 0 < nf < st.probes:
            chosen = mid
            break
        if nf == 0:
            hi = mid
        else:
            lo = mid
    assert chosen is not None
    offs = level_offsets(c)
    upto = dict(intervals=offs[level], merged=offs[level + 1])
    b = lux.Batch(rooms, surrogates={i: (s, level)})
    for bound in (0.0, chosen):
        c.gi_update_surrogate(s, level, bound, stats=False)
        c.synchronize()
        ref = state(c)
        for k in range(len(rooms)):
            poison(rooms[k])
        b.set_surrogate_bound(i, bound)
        b.set_frame(4, 4)
        b.update()
        got = batch_states(rooms, b)
        assert_same(ref, got[i], f"scene {i}, bound {bound}", upto=upto)
        for k in range(len(rooms)):
            if k != i:
                assert_same(refs[k], got[k], f"scene {k}")
        b.capture()
        poison(c)
        b.launch()
        assert_same(ref, batch_states(rooms, b)[i], f"scene {i}, bound {bound}, graph", upto=upto)
    with pytest.raises(lux.LuxError):
        b.set_surrogate_bound(0, 1.0)                               # no surrogate on scene 0
    b.close()
    s.close()


def level_offsets(ctx):
    """Start of every level in the interval array (level L has
    ceil-halved dims^3 probes x 16 x 4^L directions)."""
    p = ctx._gi
    offs = [0]
    dims = list(p.dims)
    for L in range(p.levels):
        dd = dims
        offs.append(offs[-1] + int(np.prod(dd)) * p.base_phi * p.base_y * 4 ** L)
        dims = [(x + 1) // 2 for x in dims]
    return offs


def test_gpu_batch_autotuner_keeps_results(rooms):
    refs = per_scene(rooms, 3, 1)
    b = lux.Batch(rooms)
    res = b.autotune(repeats=1)
    assert res["evaluations"] >= 10 and res["ms"] > 0
    assert b.config()["waves_per_cu"] == res["waves_per_cu"]
    for c in rooms:
        poison(c)
    b.set_frame(3, 1)
    b.update()
    for i, (r, g) in enumerate(zip(refs, batch_states(rooms, b))):
        assert_same(r, g, f"tuned, scene {i}")
    b.close()


def test_gpu_batch_query_budget_per_scene(rooms):
    """Each scene's admission budget is its own: the batch never admits more
    than max(limit / 2, 1) queries of a scene and counts every attempt."""
    c = rooms[1]
    n = c.gi_sizes()[0]
    p = lux.GiParams()
    ctypes.memmove(ctypes.byref(p), ctypes.byref(c._gi), ctypes.sizeof(p))
    p.query_limit = n
    c.gi_configure(p)
    try:
        b = lux.Batch(rooms)
        b.set_frame(5, 5)
        b.update()
        b.synchronize()
        attempts, admitted = c.gi_query_counters()
        assert admitted <= n // 2 < attempts
        iv = c.gi_download()["intervals"]
        assert (iv[:, 3] == -1.0).sum() > 0
        for j in (0, 2, 3, 4):
            a2, ad2 = rooms[j].gi_query_counters()
            assert a2 == ad2 and a2 >= rooms[j].gi_sizes()[0]
        b.close()
    finally:
        p.query_limit = 0
        c.gi_configure(p)


def test_gpu_batch_counters_past_the_lds_tally():
    """The batch walks tally unbounded admissions per workgroup in LDS for the
    first 256 scenes and use the scenes' counters directly past them: with 300
    small rooms every unbounded scene's intervals, probes and query counters
    equal its own lux_gi_update, and a bounded scene on either side of 256
    keeps its budget (never more than max(limit / 2, 1) admitted)."""
    ctxs, bounded = [], (10, 290)
    try:
        for i in range(300):
            ctxs.append(make_scene(1000 + i, (120, 3.0, 2, 2, 1, 2)))
        for j in bounded:
            p = lux.GiParams()
            ctypes.memmove(ctypes.byref(p), ctypes.byref(ctxs[j]._gi), ctypes.sizeof(p))
            p.query_limit = ctxs[j].gi_sizes()[0]
            ctxs[j].gi_configure(p)
        refs = per_scene(ctxs, 4, 2)
        for c in ctxs:
            poison(c)
        b = lux.Batch(ctxs)
        assert b.sizes()["scenes"] == 300
        b.set_frame(4, 2)
        b.update()
        got = batch_states(ctxs, b)
        b.close()
        for i, (r, g) in enumerate(zip(refs, got)):
            if i in bounded:
                attempts, admitted = g["counters"]
                assert admitted <= ctxs[i].gi_sizes()[0] // 2 < attempts, (i, attempts, admitted)
            else:
                assert_same(r, g, f"scene {i}")
    finally:
        for c in ctxs:
            c.close()


def test_gpu_batch_refresh_follows_a_reconfigured_surrogate_scene():
    """A scene with a surrogate and a finite bound is reconfigured to a larger
    grid: after lux_batch_refresh its fallback words cover the new probe count
    and the batch update equals the scene's own lux_gi_update_surrogate; a
    configuration without the surrogate's level makes the refresh fail."""
    ctxs = [make_scene(20, SPECS[0]), make_scene(24, SPECS[4])]
    c = ctxs[1]
    tri, size, d, levels, R, ns = SPECS[4]
    s = lux.Surrogate(c, lux.SurrogateParams.default((0, 0, 0), (size, size, size), seed=3))
    level, bound = 1, 0.0                                           # bound 0: every probe of the level traced
    try:
        b = lux.Batch(ctxs, surrogates={1: (s, level)})
        b.set_surrogate_bound(1, bound)
        b.set_frame(2, 2)
        b.update()
        b.synchronize()
        big = 2 * d                                                 # 8x the level's probes
        sp = size / big
        c.gi_configure(lux.gi_params((0.5 * sp, 0.5 * sp, 0.5 * sp), sp, (big, big, big), levels, base_phi=4,
                                     base_y=4, interval_length=sp, emissive_samples=ns, rays_per_interval=R,
                                     seed=4, terminal_tail=1))
        b.refresh()
        st = c.gi_update_surrogate(s, level, bound)
        assert st.fallback_probes == st.probes
        c.synchronize()
        ref = state(c)
        poison(c)
        b.set_frame(2, 2)
        c.gi_set_frame(2, 2)
        c.gi_update_surrogate(s, level, bound, stats=False)
        c.synchronize()
        ref = state(c)
        poison(c)
        b.update()
        offs = level_offsets(c)
        assert_same(ref, batch_states(ctxs, b)[1], "reconfigured scene", upto=dict(intervals=offs[level],
                                                                                   merged=offs[level + 1]))
        c.gi_configure(lux.gi_params((0.5 * sp, 0.5 * sp, 0.5 * sp), sp, (big, big, big), 1, base_phi=4,
                                     base_y=4, interval_length=sp, emissive_samples=ns, seed=4, terminal_tail=1))
        with pytest.raises(lux.LuxError):
            b.refresh()                                             # level 1 is not below 1 level
        b.set_surrogate(1, None)
        b.refresh()
        b.close()
    finally:
        s.close()
        for x in ctxs:
            x.close()
