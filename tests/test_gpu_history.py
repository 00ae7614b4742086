"""HIP temporal interval history against the oracle, bit for bit.

Cornell box, history on (the default component's blends 0.88 / 0.98, 24
frames, persistent world intervals): over four updates with changing frame
and seed -- the last one after a dirty region near the floor -- the
supported-ray counts, blended intervals, history (values and masses),
merged radiance and packed probes equal the oracle's trace, history pass,
merge and pack.  A batch with history scenes equals their own updates.
"""
import numpy as np
import pytest

from bevy_penumbra_amd import lux, scenes
from oracle import oracle as O
from tests.transport_cases import cornell_params

pytestmark = pytest.mark.gpu

RB, TB, FRAMES = 0.88, 0.98, 24


def to_lux(p):
    q = lux.GiParams()
    for name, _ in lux.GiParams._fields_:
        setattr(q, name, getattr(p, name))
    return q


def emitter_bounds(lights):
    pts = []
    for l in lights:
        a = l[0:3]
        pts += [a, a + l[12:15], a + l[16:19]]  # corner a, a + edge1, a + edge2
    pts = np.array(pts, np.float32)
    return pts.min(axis=0), pts.max(axis=0)


def bits_equal(a, b, what):
    a = np.ascontiguousarray(a).view(np.uint32)
    b = np.ascontiguousarray(b).view(np.uint32)
    bad = np.nonzero(np.any((a != b).reshape(a.shape[0], -1), axis=1))[0]
    assert bad.size == 0, (what, bad[:5])


@pytest.mark.parametrize("mode", [0, 1])
def test_gpu_history_matches_oracle_over_frames(lux_ctx, mode):
    sc, _ = scenes.cornell()
    p = cornell_params(rays_per_interval=2)
    lux_ctx.upload_scene(sc)
    lux_ctx.build_bvh()
    lux_ctx.gi_set_trace_mode(mode)
    lux_ctx.gi_configure(to_lux(p))
    lux_ctx.gi_set_history(RB, TB, FRAMES, 1)
    built = O.build_scene(sc)
    lights = O.build_lights(sc)
    emit = emitter_bounds(lights)
    n = O.total_intervals(p)
    hist = np.zeros((n, 8), np.float32)
    dirty = ((0.1, 0.1, 0.1), (0.3, 0.3, 0.3))
    try:
        for f in range(4):
            boxes = []
            if f == 3:
                lux_ctx.gi_invalidate_history(*dirty)
                boxes = [dirty]
            lux_ctx.gi_set_frame(10 + f, 7 + f)
            lux_ctx.gi_update()
            lux_ctx.synchronize()
            g = lux_ctx.gi_download()
            gh, gs = lux_ctx.gi_history_download()
            p.frame, p.seed = 10 + f, 7 + f
            iv, ns = O.gi_trace(built, lights, sc.materials, p, threads=16, samples=True)
            if f == 3:
                _, kept_all = O.gi_history(p, RB, TB, FRAMES, 1, True, iv, ns, hist, emit=emit)
            blended, hist = O.gi_history(p, RB, TB, FRAMES, 1, f > 0, iv, ns, hist, emit=emit, boxes=boxes)
            merged = O.gi_merge(p, blended, hist)
            probes = O.gi_pack(p, merged)
            assert np.array_equal(gs, ns), f
            bits_equal(g["intervals"], blended, f"intervals, update {f}")
            bits_equal(gh, hist, f"history, update {f}")
            bits_equal(g["merged"], merged, f"merged, update {f}")
            bits_equal(g["probes"], probes, f"probes, update {f}")
            if f == 1:
                assert np.all(hist[ns > 0, 3] > 1.0)       # masses grow past one update
            if f == 3:
                dropped = hist[:, 3] != kept_all[:, 3]
                assert 0 < dropped.sum() < dropped.size // 2, dropped.sum()   # the box drops some, not all
    finally:
        lux_ctx.gi_history_off()
        lux_ctx.gi_set_trace_mode(0)


def test_gpu_batch_with_history_equals_own_updates():
    sc, _ = scenes.cornell()
    ctxs = []
    try:
        for i, dims in enumerate([(16, 16, 16), (8, 8, 8)]):
            c = lux.Context(0)
            c.upload_scene(sc)
            c.build_bvh()
            p = cornell_params(seed=i)
            p.dims[:] = dims
            p.spacing = 2.0 / dims[0]
            c.gi_configure(to_lux(p))
            if i == 0:
                c.gi_set_history(RB, TB, FRAMES, 1)
            ctxs.append(c)
        refs = []
        for f in range(3):
            row = []
            for c in ctxs:
                c.gi_set_frame(f, f)
                c.gi_update()
                c.synchronize()
                row.append((c.gi_download(), c.gi_history_download()[0]))
            refs.append(row)
        for c in ctxs:
            c.gi_reset_history()
        b = lux.Batch(ctxs)
        b.capture()
        for f in range(3):
            b.set_frame(f, f)
            b.launch()
            b.synchronize()
            for i, c in enumerate(ctxs):
                g = c.gi_download()
                for key in ("intervals", "merged", "probes"):
                    bits_equal(g[key], refs[f][i][0][key], f"frame {f} scene {i} {key}")
                if i == 0:
                    bits_equal(c.gi_history_download()[0], refs[f][i][1], f"frame {f} history")
        b.close()
    finally:
        for c in ctxs:
            c.close()
