"""Scenes installed from the reference's packed words (lux_scene_upload_words).

The scene and geometry words of tests/scene_words.py (the reference's
layout over SAH trees) go through the C ABI; the device repack must give
exactly the oracle scene of that structure: hits bit for bit the oracle's
traversal, and the GI update (per-instance materials, emitter table from
the parsed geometry) bit for bit the oracle's trace, merge and pack with
that emitter table.
"""
import numpy as np
import pytest

from bevy_penumbra_amd import lux, scenes
from oracle import oracle as O
from tests import scene_words as W
from tests.gpu_util import trace_gpu

pytestmark = pytest.mark.gpu


def test_gpu_words_scene_traversal_and_gi_match_oracle(lux_ctx):
    meshes, instances, materials = W.room_scene(2)
    sw, gw, built, order = W.pack(meshes, instances, materials.shape[0])
    lux_ctx.upload_words(sw, gw, materials)
    lux_ctx.build_bvh()                      # installed built: nothing to do
    rays = scenes.microbench_rays(((-3.0, 0.0, -3.0), (3.0, 4.0, 3.0)), n=200000, seed=9)
    hits = trace_gpu(lux_ctx, rays)
    want = built.trace(rays)
    bad = np.nonzero(np.any(hits != want, axis=1))[0]
    assert bad.size == 0, (bad[:5], hits[bad[:5]], want[bad[:5]])
    assert (hits[:, 1] != 0xFFFFFFFF).mean() > 0.3
    p = lux.gi_params((-2.75, 0.25, -2.75), 0.5, (12, 8, 12), 3, base_phi=4, base_y=4, interval_length=0.5,
                      emissive_samples=8, seed=3, terminal_tail=1)
    lux_ctx.gi_configure(p)
    lux_ctx.gi_update()
    lux_ctx.synchronize()
    g = lux_ctx.gi_download()
    lights = g["lights"]
    assert lights.shape[0] == 2                      # the emissive panel's two triangles
    assert np.allclose(lights[:, 4:7], [12, 9, 4])
    po = O.GiParams()
    for name, _ in O.GiParams._fields_:
        setattr(po, name, getattr(p, name))
    iv = O.gi_trace(built, lights, materials, po, threads=16)
    for key, val in (("intervals", iv), ("merged", O.gi_merge(po, iv))):
        assert np.array_equal(g[key].view(np.uint32), val.view(np.uint32)), key
    assert np.array_equal(g["probes"].view(np.uint32), O.gi_pack(po, O.gi_merge(po, iv)).view(np.uint32))
    with pytest.raises(lux.LuxError):
        lux_ctx.set_transforms(np.stack([i["inv"] for i in instances]), np.stack([i["fwd"] for i in instances]))


def test_gpu_malformed_words_and_records_are_refused(lux_ctx):
    meshes, instances, materials = W.room_scene(1)
    sw, gw, built, _ = W.pack(meshes, instances, materials.shape[0])
    bad = sw.copy()
    bad[0, 3] = 8                                    # instance stride other than 9 words
    with pytest.raises(lux.LuxError):
        lux_ctx.upload_words(bad, gw, materials)
    bad = sw.copy()
    base = int(sw[0, 1])
    bad[base + 8, 2] = gw[0, 1]                      # triangle count past the index words
    with pytest.raises(lux.LuxError):
        lux_ctx.upload_words(bad, gw, materials)
    bad_rec = built.inst.copy()
    bad_rec[0, 27] = built.tri.shape[0] + 1          # instance range past the triangle table
    with pytest.raises(lux.LuxError):
        lux_ctx.upload_built(built.blas, built.tlas, built.tri, bad_rec)
