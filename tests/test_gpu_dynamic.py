"""Dynamic frame pieces against the oracle: per-frame rigid motion.

* the TLAS rebuilt after lux_scene_set_transforms (Morton-30 over instance
  boxes) is bit-identical to the oracle's build of the moved scene, and rays
  through it return the oracle's hits;
* a forced full rebuild (every BLAS too) reproduces the same nodes;
* the emitter table follows the instances (build_emissive_lights on the
  moved scene, src/scene.rs:2311-2372);
* camera receivers are the oracle's closest hits of the same pinhole rays.
"""
import dataclasses

import numpy as np
import pytest

from bevy_penumbra_amd import lux, scenes
from bevy_penumbra_amd.lux import DeviceBuffer
from oracle import oracle as O
from tests.gpu_util import trace_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_dynamic():
    # 32 instances of 32 distinct 256-triangle meshes
    return scenes.dynamic(seed=3, meshes=32, instances=32, stacks=8, slices=16)


def moved(sc, motion, frame):
    inv, fwd = scenes.dynamic_transforms(motion, frame)
    return dataclasses.replace(sc, inst_inv=inv, inst_fwd=fwd), inv, fwd


def test_gpu_tlas_rebuild_follows_motion_bit_exact(lux_ctx, small_dynamic):
    sc, motion = small_dynamic
    lux_ctx.upload_scene(sc)
    lux_ctx.build_bvh()
    for frame in (0, 7, 60, 119):
        sc2, inv, fwd = moved(sc, motion, frame)
        lux_ctx.set_transforms(inv, fwd)
        lux_ctx.build_bvh()
        built = O.build_scene(sc2)
        g = lux_ctx.download_bvh()
        assert np.array_equal(g["blas"], built.blas), frame
        assert np.array_equal(g["tlas"], built.tlas), frame
        assert np.array_equal(g["inst_order"], built.inst[:, 28]), frame
        rays = scenes.microbench_rays(sc.bounds, n=65536, seed=frame + 1)
        hits = trace_gpu(lux_ctx, rays)
        want = built.trace(rays)
        bad = np.nonzero(np.any(hits != want, axis=1))[0]
        assert bad.size == 0, (frame, bad[:5])


def test_gpu_forced_full_rebuild_is_identical(lux_ctx, small_dynamic):
    sc, motion = small_dynamic
    sc2, inv, fwd = moved(sc, motion, 33)
    lux_ctx.upload_scene(sc)
    lux_ctx.build_bvh()
    lux_ctx.set_transforms(inv, fwd)
    lux_ctx.build_bvh()
    a = lux_ctx.download_bvh()
    lux_ctx.build_bvh(lux.LUX_BVH_MORTON30 | lux.LUX_BVH_FORCE_BLAS)
    b = lux_ctx.download_bvh()
    for k in ("blas", "tlas", "tri_order", "inst_order"):
        assert np.array_equal(a[k], b[k]), k


def test_gpu_emitter_table_follows_motion(lux_ctx, small_dynamic):
    sc, motion = small_dynamic
    lux_ctx.upload_scene(sc)
    lux_ctx.build_bvh()
    lux_ctx.gi_configure(lux.gi_params((-0.5, -0.5, -0.5), 2.0, (16, 6, 8), 2))
    for frame in (0, 45):
        sc2, inv, fwd = moved(sc, motion, frame)
        lux_ctx.set_transforms(inv, fwd)
        lux_ctx.build_bvh()
        lux_ctx.gi_update()
        lights = lux_ctx.gi_download()["lights"]
        want = O.build_lights(sc2)
        assert lights.shape == want.shape and lights.shape[0] == 32 * 32
        assert np.array_equal(lights.view(np.uint32), want.view(np.uint32)), frame


def camera_rays_f32(cam, w, h):
    """The pinhole rays of lux_camera_receivers in the kernel's f32 order."""
    f32 = np.float32
    o, fw, rt, up = (cam[3 * k:3 * k + 3].astype(f32) for k in range(4))
    x = np.tile(np.arange(w, dtype=f32), h)
    y = np.repeat(np.arange(h, dtype=f32), w)
    u = (x + f32(0.5)) / f32(w) * f32(2.0) - f32(1.0)
    v = f32(1.0) - (y + f32(0.5)) / f32(h) * f32(2.0)
    d = np.stack([(fw[a] + u * rt[a]) + v * up[a] for a in range(3)], axis=1).astype(f32)
    ln = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    d = (d / ln[:, None]).astype(f32)
    rays = np.zeros((w * h, 8), f32)
    rays[:, 0:3] = o
    rays[:, 3] = f32(1e20)
    rays[:, 4:7] = d
    return rays


def test_gpu_camera_receivers_are_oracle_hits(lux_ctx, small_dynamic):
    sc, motion = small_dynamic
    sc2, inv, fwd = moved(sc, motion, 12)
    lux_ctx.upload_scene(sc2)
    lux_ctx.build_bvh()
    w, h = 320, 180
    cam = scenes.camera_rays((15.0, 6.0, -9.0), (15.0, 6.0, 7.5), w, h)
    out = DeviceBuffer.empty(lux_ctx, (w * h, 6), np.float32)
    lux_ctx.camera_receivers(cam, w, h, out)
    lux_ctx.synchronize()
    r = out.numpy()
    rays = camera_rays_f32(cam, w, h)
    hits = O.build_scene(sc2).trace(rays)
    hit = hits[:, 1] != lux.LUX_EMPTY
    assert 0.05 < hit.mean() < 0.95
    t = hits[:, 0].view(np.float32)
    pos = (rays[:, 0:3] + rays[:, 4:7] * t[:, None]).astype(np.float32)
    assert np.array_equal(r[hit, :3].view(np.uint32), pos[hit].view(np.uint32))
    assert np.all(r[~hit, 3:6] == 0.0)
    n = r[hit, 3:6]
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-5)
    assert np.all(np.sum(n * rays[hit, 4:7], axis=1) <= 0.0)
    # perpendicular to the world-space triangle that was hit
    F = sc2.inst_fwd.reshape(-1, 3, 4).astype(np.float64)
    inst = hits[hit, 2]
    tri = hits[hit, 1]
    v = np.stack([sc2.meshes[sc2.inst_mesh[i]][k].reshape(3, 3) for i, k in zip(inst, tri)])
    wv = np.einsum("nij,nkj->nki", F[inst][:, :, :3], v) + F[inst][:, None, :, 3]
    e1, e2 = wv[:, 1] - wv[:, 0], wv[:, 2] - wv[:, 0]
    for e in (e1, e2):
        cos = np.abs(np.sum(n * e, axis=1)) / np.linalg.norm(e, axis=1)
        assert cos.max() < 1e-4


def test_gpu_tlas_refit_matches_oracle_refit(lux_ctx, small_dynamic):
    """LUX_BVH_REFIT after motion keeps the TLAS topology and instance order
    and refits every box (src/scene.rs:1040-1058): bit for bit the oracle's
    refit of the pre-motion TLAS with the moved instances' bounds, and the
    oracle's traversal of that refitted tree gives the GPU's hits.  A refit
    without motion reproduces the built TLAS."""
    sc, motion = small_dynamic
    sc0, inv0, fwd0 = moved(sc, motion, 0)
    lux_ctx.upload_scene(sc0)
    lux_ctx.build_bvh()
    g0 = lux_ctx.download_bvh()
    lux_ctx.set_transforms(inv0, fwd0)
    lux_ctx.build_bvh(lux.LUX_BVH_MORTON30 | lux.LUX_BVH_REFIT)
    same = lux_ctx.download_bvh()
    assert np.array_equal(same["tlas"], g0["tlas"]) and np.array_equal(same["inst_order"], g0["inst_order"])
    order = g0["inst_order"]
    tlas = g0["tlas"]
    for frame in (5, 40):
        sc2, inv, fwd = moved(sc, motion, frame)
        lux_ctx.set_transforms(inv, fwd)
        lux_ctx.build_bvh(lux.LUX_BVH_MORTON30 | lux.LUX_BVH_REFIT)
        g = lux_ctx.download_bvh()
        built = O.build_scene(sc2)
        roots = {int(r[28]): built.blas[int(r[25])] for r in built.inst}
        lo = np.zeros((len(order), 4), np.float32)
        hi = np.zeros((len(order), 4), np.float32)
        for c in range(len(order)):
            lo[c], hi[c] = O.instance_bounds(fwd[c], roots[c])
        want = O.tlas_refit(tlas, lo, hi, order)
        assert np.array_equal(g["inst_order"], order), frame
        assert np.array_equal(g["tlas"], want), frame
        assert not np.array_equal(want, tlas)
        by_caller = {int(r[28]): r for r in built.inst}
        inst = np.stack([by_caller[int(c)] for c in order])
        refit_scene = O.BuiltScene(built.blas, want, built.tri, inst)
        rays = scenes.microbench_rays(sc.bounds, n=65536, seed=frame + 11)
        hits = trace_gpu(lux_ctx, rays)
        expect = refit_scene.trace(rays)
        bad = np.nonzero(np.any(hits != expect, axis=1))[0]
        assert bad.size == 0, (frame, bad[:5])
        tlas = want
