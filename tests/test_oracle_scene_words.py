"""The restated packed-scene format against the oracle's own scene build.

The scene and geometry words written by tests/scene_words.py (the
reference's geometry_words / node_words / pack_instances layout over SAH
trees) describe the same world as the scene given directly: the oracle's
traversal of the words' structure finds, for every ray, the same closest
distance and instance as the oracle's traversal of its own SAH build of the
same triangles and transforms, and the packed words hold what the layout
says (header, node words, instance records, geometry header, indices,
positions).
"""
import numpy as np

from bevy_penumbra_amd import scenes
from oracle import oracle as O
from tests import scene_words as W


def test_packed_words_layout_and_world():
    meshes, instances, materials = W.room_scene(1)
    sw, gw, built, order = W.pack(meshes, instances, materials.shape[0])
    n_inst = len(instances)
    end, base, count, stride = sw[0]
    assert (count, stride) == (n_inst, 9) and end == base + 9 * n_inst + 8 * materials.shape[0]
    nodes = sw[1:base].reshape(-1, 8)
    assert np.array_equal(nodes, built.tlas)
    assert gw[0, 0] == 1 and gw[0, 2] == 5 and gw[0, 1] % 5 == 0
    for slot, k in enumerate(order):
        w = sw[base + 9 * slot: base + 9 * (slot + 1)]
        assert np.array_equal(w[3:6].reshape(-1).view(np.float32), instances[k]["fwd"])
        assert w[7, 3] == end - 8 * materials.shape[0] + 8 * instances[k]["material"]
        root, first, cnt = w[8, :3]
        assert 0 < root < nodes.shape[0] and cnt == meshes[instances[k]["mesh"]].shape[0]
        t0 = int(first)
        idx = gw[1 + t0, :3]
        pos = gw[gw[0, 1] + 5 * idx, :3].view(np.float32)
        assert np.array_equal(pos.reshape(-1), built.tri[t0, [0, 1, 2, 4, 5, 6, 8, 9, 10]])
    # the same world, given directly, traced by the oracle's own SAH build
    sc = scenes.Scene(meshes=[np.ascontiguousarray(m, np.float32) for m in meshes],
                      mesh_materials=[np.zeros(m.shape[0], np.uint32) for m in meshes],
                      inst_inv=np.stack([i["inv"] for i in instances]), inst_fwd=np.stack([i["fwd"] for i in instances]),
                      inst_mesh=np.array([i["mesh"] for i in instances], np.uint32),
                      inst_flags=np.array([i["flags"] for i in instances], np.uint32), materials=materials,
                      bounds=((-3, 0, -3), (3, 4, 3)))
    direct = O.build_scene(sc, sah=True)
    rays = scenes.microbench_rays(sc.bounds, n=20000, seed=4)
    a = built.trace(rays)
    b = direct.trace(rays)
    # words hits report the word slot: map to the caller's instance
    inst_a = np.where(a[:, 2] == 0xFFFFFFFF, 0xFFFFFFFF, order[np.minimum(a[:, 2], len(order) - 1)])
    assert np.array_equal(a[:, 0], b[:, 0])            # closest distance (or tmax) bits
    assert np.array_equal(inst_a, b[:, 2])
    assert (a[:, 1] != 0xFFFFFFFF).mean() > 0.5
