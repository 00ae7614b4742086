"""ctypes loader for the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, the smoke check and bench.py's cpu_baseline leg import this
module; the product path (bevy_penumbra_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liblux_oracle.so"
_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    lib = C.CDLL(str(LIB))
    vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int
    sig = {
        "orc_box_interval": (i32, [vp, vp, vp, vp, C.c_float, vp, vp]),
        "orc_intersect_triangle": (None, [vp, vp, vp, vp, vp, vp]),
        "orc_trace": (None, [vp, vp, u32, vp, vp, vp, i32]),
        "orc_lbvh_build": (u32, [vp, vp, u32, u32, vp, vp, vp]),
        "orc_morton_codes": (None, [vp, vp, u32, vp]),
        "orc_stable_sort_pairs": (None, [vp, u32, vp]),
        "orc_scene_build": (i32, [vp, vp, vp, u32, vp, vp, vp, vp, u32, u32, i32, vp, vp, vp, vp, vp]),
        "orc_scene_capacity": (None, [vp, u32, u32, vp, vp, vp]),
        "orc_sah_build": (u32, [vp, vp, u32, vp, vp]),
        "orc_tree_cost": (C.c_double, [vp, u32]),
        "orc_tree_depth": (u32, [vp, u32]),
        "orc_brute_force": (u32, [vp, vp, vp, vp, vp, u32]),
        "orc_linked_sort": (u32, [vp, u32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data


class OrcScene(C.Structure):
    _fields_ = [("tlas", C.c_void_p), ("n_tlas", C.c_uint32), ("inst", C.c_void_p),
                ("n_inst", C.c_uint32), ("blas", C.c_void_p), ("n_blas", C.c_uint32),
                ("tri", C.c_void_p), ("n_tri", C.c_uint32)]


class OrcStats(C.Structure):
    _fields_ = [("node_fetches", C.c_uint64), ("tri_tests", C.c_uint64), ("inst_visits", C.c_uint64)]


class BuiltScene:
    """Oracle-built scene: arrays in the same layout the GPU builds."""

    def __init__(self, blas, tlas, tri, inst):
        self.blas, self.tlas, self.tri, self.inst = blas, tlas, tri, inst
        self._s = OrcScene(tlas.ctypes.data, tlas.shape[0], inst.ctypes.data, inst.shape[0],
                           blas.ctypes.data, blas.shape[0], tri.ctypes.data, tri.shape[0])

    def trace(self, rays: np.ndarray, threads: int = 8, stats: bool = False, bary: bool = False):
        lib = load()
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        hits = np.zeros((n, 4), np.uint32)
        b = np.zeros((n, 2), np.float32) if bary else None
        st = OrcStats() if stats else None
        lib.orc_trace(C.byref(self._s), rays.ctypes.data, n, hits.ctypes.data, _p(b),
                      C.byref(st) if st is not None else None, threads)
        out = [hits]
        if bary:
            out.append(b)
        if stats:
            out.append(dict(node_fetches=st.node_fetches, tri_tests=st.tri_tests,
                            inst_visits=st.inst_visits))
        return out[0] if len(out) == 1 else tuple(out)

    def brute_force(self, ray: np.ndarray, cap: int = 64):
        lib = load()
        ray = np.ascontiguousarray(ray, np.float32)
        ts = np.zeros(cap, np.float64)
        tris = np.zeros(cap, np.uint32)
        insts = np.zeros(cap, np.uint32)
        k = lib.orc_brute_force(C.byref(self._s), ray.ctypes.data, ts.ctypes.data,
                                tris.ctypes.data, insts.ctypes.data, cap)
        return ts[:k], tris[:k], insts[:k]


def build_scene(scene, sah: bool = False, leaf_max: int = 4) -> BuiltScene:
    lib = load()
    nm = len(scene.meshes)
    pos = [np.ascontiguousarray(m, np.float32) for m in scene.meshes]
    mat = [np.ascontiguousarray(m, np.uint32) for m in scene.mesh_materials]
    counts = np.asarray([p.shape[0] for p in pos], np.uint32)
    pos_ptrs = (C.c_void_p * nm)(*[p.ctypes.data for p in pos])
    mat_ptrs = (C.c_void_p * nm)(*[m.ctypes.data for m in mat])
    ni = scene.inst_mesh.shape[0]
    tri_cap, blas_cap, tlas_cap = C.c_uint32(), C.c_uint32(), C.c_uint32()
    lib.orc_scene_capacity(counts.ctypes.data, nm, ni, C.byref(tri_cap), C.byref(blas_cap),
                           C.byref(tlas_cap))
    tri = np.zeros((tri_cap.value, 12), np.float32)
    blas = np.zeros((blas_cap.value, 8), np.uint32)
    tlas = np.zeros((tlas_cap.value, 8), np.uint32)
    inst = np.zeros((max(1, ni), 32), np.uint32)
    sizes = np.zeros(4, np.uint32)
    inv = np.ascontiguousarray(scene.inst_inv, np.float32)
    fwd = np.ascontiguousarray(scene.inst_fwd, np.float32)
    im = np.ascontiguousarray(scene.inst_mesh, np.uint32)
    fl = np.ascontiguousarray(scene.inst_flags, np.uint32)
    lib.orc_scene_build(pos_ptrs, mat_ptrs, counts.ctypes.data, nm, inv.ctypes.data,
                        fwd.ctypes.data, im.ctypes.data, fl.ctypes.data, ni, leaf_max,
                        1 if sah else 0, tri.ctypes.data, blas.ctypes.data, tlas.ctypes.data,
                        inst.ctypes.data, sizes.ctypes.data)
    return BuiltScene(blas[: sizes[1]].copy(), tlas[: sizes[2]].copy(), tri[: sizes[0]].copy(),
                      inst[:ni].copy())


def morton_codes(lo4: np.ndarray, hi4: np.ndarray) -> np.ndarray:
    lo4 = np.ascontiguousarray(lo4, np.float32)
    hi4 = np.ascontiguousarray(hi4, np.float32)
    out = np.zeros(lo4.shape[0], np.uint32)
    load().orc_morton_codes(lo4.ctypes.data, hi4.ctypes.data, lo4.shape[0], out.ctypes.data)
    return out


def lbvh(lo4: np.ndarray, hi4: np.ndarray, leaf_max: int = 4):
    n = lo4.shape[0]
    lo4 = np.ascontiguousarray(lo4, np.float32)
    hi4 = np.ascontiguousarray(hi4, np.float32)
    codes = np.zeros(n, np.uint32)
    perm = np.zeros(n, np.uint32)
    nodes = np.zeros((max(1, 2 * n - 1), 8), np.uint32)
    k = load().orc_lbvh_build(lo4.ctypes.data, hi4.ctypes.data, n, leaf_max, codes.ctypes.data,
                              perm.ctypes.data, nodes.ctypes.data)
    return codes, perm, nodes[:k]


def sah(lo4: np.ndarray, hi4: np.ndarray):
    n = lo4.shape[0]
    lo4 = np.ascontiguousarray(lo4, np.float32)
    hi4 = np.ascontiguousarray(hi4, np.float32)
    perm = np.zeros(n, np.uint32)
    nodes = np.zeros((max(1, 2 * n - 1), 8), np.uint32)
    k = load().orc_sah_build(lo4.ctypes.data, hi4.ctypes.data, n, perm.ctypes.data, nodes.ctypes.data)
    return perm, nodes[:k]


def tree_cost(nodes: np.ndarray) -> float:
    nodes = np.ascontiguousarray(nodes)
    return load().orc_tree_cost(nodes.ctypes.data, nodes.shape[0])


def tree_depth(nodes: np.ndarray) -> int:
    nodes = np.ascontiguousarray(nodes)
    return load().orc_tree_depth(nodes.ctypes.data, nodes.shape[0])


def stable_sort_pairs(keys: np.ndarray) -> np.ndarray:
    keys = np.ascontiguousarray(keys, np.uint32)
    perm = np.zeros(keys.shape[0], np.uint32)
    load().orc_stable_sort_pairs(keys.ctypes.data, keys.shape[0], perm.ctypes.data)
    return perm


def linked_sort(nodes: np.ndarray, head: int) -> int:
    assert nodes.dtype == np.uint32 and nodes.flags["C_CONTIGUOUS"]
    return load().orc_linked_sort(nodes.ctypes.data, head)


def intersect_triangle(o, d, a, b, c):
    arrs = [np.ascontiguousarray(x, np.float32) for x in (o, d, a, b, c)]
    out = np.zeros(3, np.float32)
    load().orc_intersect_triangle(*[x.ctypes.data for x in arrs], out.ctypes.data)
    return out


def box_interval(o, d, mn, mx, limit):
    arrs = [np.ascontiguousarray(x, np.float32) for x in (o, d, mn, mx)]
    e, x = C.c_float(), C.c_float()
    hit = load().orc_box_interval(*[a.ctypes.data for a in arrs], float(limit), C.byref(e), C.byref(x))
    return bool(hit), e.value, x.value
