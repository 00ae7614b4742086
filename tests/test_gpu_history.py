"""HIP temporal interval history against the oracle, bit for bit.

Cornell box, history on (the default component's blends 0.88 / 0.98, 24
frames, persistent world intervals): over four updates with changing frame
and seed -- the last one after a dirty region near the floor -- the
supported-ray counts, blended intervals, history (values, masses and each
interval's retained dependency: the union of the segments its rays and NEE
occlusion rays traversed), merged radiance and packed probes equal the
oracle's trace, history pass, merge and pack.  A batch with history scenes equals their own updates.
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
    n = O.total_intervals(p)
    hist = np.zeros((n, 16), np.float32)
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
            iv, ns, dep = O.gi_trace(built, lights, sc.materials, p, threads=16, dep=True)
            if f == 3:
                _, kept_all = O.gi_history(p, RB, TB, FRAMES, 1, True, iv, ns, hist, dep)
            blended, hist = O.gi_history(p, RB, TB, FRAMES, 1, f > 0, iv, ns, hist, dep, boxes=boxes)
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
                # the box drops the intervals whose segments (or NEE segments)
                # passed through it: some, a small share
                dropped = hist[:, 3] != kept_all[:, 3]
                assert 0 < dropped.sum() < dropped.size // 4, dropped.sum()
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


def world_box(mesh_tris, fwd_row):
    """World box of a mesh's local bounds under 3x4 rows (f32, the library's
    operation order)."""
    p = mesh_tris.reshape(-1, 3)
    lo, hi = p.min(axis=0), p.max(axis=0)
    f = np.asarray(fwd_row, np.float32).reshape(3, 4)
    out_lo = np.full(3, np.inf, np.float32)
    out_hi = np.full(3, -np.inf, np.float32)
    for c in range(8):
        pl = np.array([hi[0] if c & 1 else lo[0], hi[1] if c & 2 else lo[1], hi[2] if c & 4 else lo[2]], np.float32)
        w = ((f[:, 0] * pl[0] + f[:, 1] * pl[1]) + f[:, 2] * pl[2]) + f[:, 3]
        out_lo, out_hi = np.minimum(out_lo, w), np.maximum(out_hi, w)
    return out_lo, out_hi


def test_gpu_history_after_rigid_motion_matches_oracle(lux_ctx):
    """Six meshes, every other one with an emissive band; after one update a
    non-emissive instance moves: the library marks its old and new boxes
    dirty and the next update's history (with every interval's retained
    dependency) equals the oracle's history pass over the same boxes; most
    intervals keep their history.  Moving an emissive instance drops the whole
    history."""
    sc, motion = scenes.dynamic(seed=3, meshes=6, instances=6, stacks=6, slices=8, emissive_every=2)
    p = O.gi_params((0.5, 0.5, 0.5), 1.0, (29, 11, 14), 3, base_phi=4, base_y=4, interval_mode=0,
                    interval_length=1.0, emissive_samples=8, frame=0, seed=0, terminal_tail=1)
    lux_ctx.upload_scene(sc)
    lux_ctx.build_bvh()
    lux_ctx.gi_configure(to_lux(p))
    lux_ctx.gi_set_history(RB, TB, FRAMES, 1)
    n = O.total_intervals(p)
    hist = np.zeros((n, 16), np.float32)
    try:
        lux_ctx.gi_update()
        lux_ctx.synchronize()
        built = O.build_scene(sc)
        lights = O.build_lights(sc)
        iv, ns, dep = O.gi_trace(built, lights, sc.materials, p, threads=16, dep=True)
        _, hist = O.gi_history(p, RB, TB, FRAMES, 1, False, iv, ns, hist, dep)
        bits_equal(lux_ctx.gi_history_download()[0], hist, "history after the first update")
        # one instance moves by a small step that keeps it inside the scene's
        # previous root box (a box leaving it would reset the whole history)
        inv, fwd = sc.inst_inv.copy(), sc.inst_fwd.copy()
        boxes = [world_box(sc.meshes[int(sc.inst_mesh[i])], fwd[i]) for i in range(len(fwd))]
        rlo = np.min([b[0] for b in boxes], axis=0) + 1e-3
        rhi = np.max([b[1] for b in boxes], axis=0) - 1e-3
        step = np.float32([0.25, 0.125, 0.0])
        k = next(i for i, (lo, hi) in enumerate(boxes)
                 if sc.inst_mesh[i] % 2 == 1 and np.all(lo + step >= rlo) and np.all(hi + step <= rhi))
        m = int(sc.inst_mesh[k])
        old_box = boxes[k]
        fwd[k, 3] += step[0]
        fwd[k, 7] += step[1]
        inv_m = np.linalg.inv(np.vstack([fwd[k].reshape(3, 4).astype(np.float64), [0, 0, 0, 1]]))
        inv[k] = inv_m[:3].reshape(12).astype(np.float32)
        new_box = world_box(sc.meshes[m], fwd[k])
        lux_ctx.set_transforms(inv, fwd)
        lux_ctx.build_bvh()
        lux_ctx.gi_set_frame(1, 1)
        lux_ctx.gi_update()
        lux_ctx.synchronize()
        reused = lux_ctx.gi_history_reused()
        g = lux_ctx.gi_download()
        gh, gs = lux_ctx.gi_history_download()
        sc.inst_inv, sc.inst_fwd = inv, fwd
        built = O.build_scene(sc)
        lights = O.build_lights(sc)
        p.frame, p.seed = 1, 1
        iv, ns, dep = O.gi_trace(built, lights, sc.materials, p, threads=16, dep=True)
        blended, hist = O.gi_history(p, RB, TB, FRAMES, 1, True, iv, ns, hist, dep,
                                     boxes=[old_box, new_box])
        assert np.array_equal(gs, ns)
        bits_equal(g["intervals"], blended, "intervals after the motion")
        bits_equal(gh, hist, "history after the motion")
        merged = O.gi_merge(p, blended, hist)
        bits_equal(g["merged"], merged, "merged after the motion")
        bits_equal(g["probes"], O.gi_pack(p, merged), "probes after the motion")
        assert 0.5 * n < reused < n, (reused, n)
        e = next(i for i in range(len(fwd)) if sc.inst_mesh[i] % 2 == 0)   # an emitting instance
        fwd2, inv2 = fwd.copy(), inv.copy()
        fwd2[e, 7] += np.float32(0.0625)
        inv_m = np.linalg.inv(np.vstack([fwd2[e].reshape(3, 4).astype(np.float64), [0, 0, 0, 1]]))
        inv2[e] = inv_m[:3].reshape(12).astype(np.float32)
        lux_ctx.set_transforms(inv2, fwd2)
        lux_ctx.build_bvh()
        lux_ctx.gi_set_frame(2, 2)
        lux_ctx.gi_update()
        assert lux_ctx.gi_history_reused() == 0
    finally:
        lux_ctx.gi_history_off()
