"""ctypes loader for the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, the smoke check and bench.py's cpu_baseline leg import this
module; the product path (bevy_penumbra_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liblux_oracle.so"
_lib = None


def load() -> C.CDLL:
    """The oracle library: oracle/liblux_oracle.so, or the build named by the
    LUX_ORACLE_LIB environment variable (the CPU baseline's host-tuned build)."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ["LUX_ORACLE_LIB"]) if os.environ.get("LUX_ORACLE_LIB") else LIB
    if not path.exists():
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    lib = C.CDLL(str(path))
    vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int
    sig = {
        "orc_box_interval": (i32, [vp, vp, vp, vp, C.c_float, vp, vp]),
        "orc_intersect_triangle": (None, [vp, vp, vp, vp, vp, vp]),
        "orc_trace": (None, [vp, vp, u32, vp, vp, vp, i32]),
        "orc_lbvh_build": (u32, [vp, vp, u32, u32, vp, vp, vp]),
        "orc_morton_codes": (None, [vp, vp, u32, vp]),
        "orc_stable_sort_pairs": (None, [vp, u32, vp]),
        "orc_scene_build": (i32, [vp, vp, vp, u32, vp, vp, vp, vp, u32, u32, i32, vp, vp, vp, vp, vp]),
        "orc_instance_bounds": (None, [vp, vp, vp, vp]),
        "orc_tlas_refit": (None, [vp, u32, vp, vp, vp, u32]),
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


def build_scene(scene, sah: int = 0, leaf_max: int = 4) -> BuiltScene:
    """sah 0: LBVH BLASes and TLAS; 1: binned-SAH BLASes (packed) and TLAS;
    2: binned-SAH BLASes in LBVH-sized regions with an LBVH TLAS (the
    library's LUX_BVH_SAH layout)."""
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
                        int(sah), tri.ctypes.data, blas.ctypes.data, tlas.ctypes.data,
                        inst.ctypes.data, sizes.ctypes.data)
    return BuiltScene(blas[: sizes[1]].copy(), tlas[: sizes[2]].copy(), tri[: sizes[0]].copy(),
                      inst[:ni].copy())


def instance_bounds(fwd_rows: np.ndarray, blas_root: np.ndarray):
    """(lo, hi) float32[4] world bounds of an instance (12 forward-row floats,
    its BLAS root node as 8 uint32 words)."""
    f = np.ascontiguousarray(fwd_rows, np.float32)
    r = np.ascontiguousarray(blas_root, np.uint32)
    lo, hi = np.zeros(4, np.float32), np.zeros(4, np.float32)
    load().orc_instance_bounds(f.ctypes.data, r.ctypes.data, lo.ctypes.data, hi.ctypes.data)
    return lo, hi


def tlas_refit(tlas: np.ndarray, lo: np.ndarray, hi: np.ndarray, order: np.ndarray) -> np.ndarray:
    """Refitted copy of TLAS nodes (n, 8) uint32 for instance boxes lo/hi
    (caller order, float32[n_inst, 4]) and the TLAS instance order."""
    t = np.array(tlas, np.uint32, copy=True, order="C")
    l = np.ascontiguousarray(lo, np.float32)
    h = np.ascontiguousarray(hi, np.float32)
    o = np.ascontiguousarray(order, np.uint32)
    load().orc_tlas_refit(t.ctypes.data, t.shape[0], l.ctypes.data, h.ctypes.data, o.ctypes.data, o.shape[0])
    return t


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


# ---------------------------------------------------------------- transport
class GiParams(C.Structure):
    """Mirror of orc_gi_params / lux_gi_params."""
    _fields_ = [("origin", C.c_float * 3), ("spacing", C.c_float), ("dims", C.c_uint32 * 3),
                ("levels", C.c_uint32), ("base_phi", C.c_uint32), ("base_y", C.c_uint32),
                ("interval_mode", C.c_uint32), ("interval_length", C.c_float),
                ("normal_bias", C.c_float), ("emissive_samples", C.c_uint32),
                ("frame", C.c_uint32), ("seed", C.c_uint32), ("rays_per_interval", C.c_uint32),
                ("terminal_tail", C.c_uint32), ("flags", C.c_uint32), ("query_limit", C.c_uint32)]


def gi_params(origin, spacing, dims, levels, base_phi=8, base_y=4, interval_mode=0,
              interval_length=0.4, normal_bias=0.005, emissive_samples=16, frame=0, seed=0,
              rays_per_interval=1, terminal_tail=1, query_limit=0) -> GiParams:
    p = GiParams()
    p.origin[:] = [float(x) for x in origin]
    p.spacing = spacing
    p.dims[:] = [int(x) for x in dims]
    p.levels = levels
    p.base_phi, p.base_y = base_phi, base_y
    p.interval_mode = interval_mode
    p.interval_length = interval_length
    p.normal_bias = normal_bias
    p.emissive_samples = emissive_samples
    p.frame, p.seed = frame, seed
    p.rays_per_interval = rays_per_interval
    p.terminal_tail = terminal_tail
    p.query_limit = query_limit
    return p


def _tsig():
    lib = load()
    vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int
    for name, res, args in [
        ("orc_gi_level_dims", u32, [vp, u32, vp]),
        ("orc_gi_level_offset", C.c_uint64, [vp, u32]),
        ("orc_gi_level_directions", u32, [vp, u32]),
        ("orc_gi_interval_bounds", None, [vp, u32, vp, vp]),
        ("orc_build_lights", u32, [vp, vp, vp, u32, vp, vp, vp, u32, vp, u32, vp, u32]),
        ("orc_bin_direction", None, [u32, u32, u32, C.c_float, C.c_float, vp]),
        ("orc_cos_sin_2pi", None, [C.c_float, vp, vp]),
        ("orc_gi_trace", None, [vp, vp, u32, vp, u32, vp, vp, i32]),
        ("orc_gi_trace_stride", None, [vp, vp, u32, vp, u32, vp, vp, i32, C.c_uint64]),
        ("orc_gi_trace_f64", None, [vp, vp, u32, vp, u32, vp, vp, i32]),
        ("orc_gi_merge", None, [vp, vp, vp]),
        ("orc_gi_trace_samples", None, [vp, vp, u32, vp, u32, vp, vp, vp, i32]),
        ("orc_gi_history", None, [vp, vp, vp, u32, C.c_uint64, vp, vp, vp, vp]),
        ("orc_gi_trace_dep", None, [vp, vp, u32, vp, u32, vp, vp, vp, vp, i32]),
        ("orc_gi_merge_history", None, [vp, vp, vp, vp]),
        ("orc_gi_merge_f64", None, [vp, vp, vp]),
        ("orc_gi_pack", None, [vp, vp, vp]),
        ("orc_gi_pack_f64", None, [vp, vp, vp]),
        ("orc_world_resolve", None, [vp, vp, vp, u32, vp]),
        ("orc_world_resolve_f64", None, [vp, vp, vp, u32, vp]),
        ("orc_lod_select", None, [vp, vp, vp, vp, vp]),
        ("orc_world_resolve_lods", None, [vp, vp, u32, vp, vp, u32, vp]),
        ("orc_world_resolve_lods_f64", None, [vp, vp, u32, vp, vp, u32, vp]),
        ("orc_gi_level_queries", None, [vp, u32, vp]),
        ("orc_gi_sample_index", C.c_uint64, [u32, u32, C.c_uint64]),
        ("orc_candidate_rank", u32, [u32, u32, u32, u32]),
        ("orc_candidate_from_rank", u32, [u32, u32, u32, u32]),
        ("orc_candidate_selected", i32, [u32, vp]),
        ("orc_rx_cell", None, [vp, C.c_float, u32, u32, vp]),
        ("orc_rx_directory_capacity", u32, [u32]),
        ("orc_rx_seed", None, [vp, vp, vp, C.c_float, u32, i32, vp, vp, vp]),
        ("orc_rx_select", u32, [u32, u32, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp]),
        ("orc_cold_candidate_rank", u32, [u32, u32, u32, u32, u32]),
        ("orc_cold_candidate_from_rank", u32, [u32, u32, u32, u32, u32]),
        ("orc_direction_index", u32, [vp, u32, u32]),
        ("orc_rx_trace", None, [vp, vp, u32, vp, u32, vp, u32, vp, u32, vp, vp, vp, u32, u32, vp, vp, vp]),
    ]:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def level_dims(p: GiParams, level: int):
    d = (C.c_uint32 * 3)()
    n = _tsig().orc_gi_level_dims(C.byref(p), level, d)
    return n, tuple(d)


def level_offset(p: GiParams, level: int) -> int:
    return int(_tsig().orc_gi_level_offset(C.byref(p), level))


def level_directions(p: GiParams, level: int) -> int:
    return int(_tsig().orc_gi_level_directions(C.byref(p), level))


def interval_bounds(p: GiParams, level: int):
    s, e = C.c_float(), C.c_float()
    _tsig().orc_gi_interval_bounds(C.byref(p), level, C.byref(s), C.byref(e))
    return s.value, e.value


def total_intervals(p: GiParams) -> int:
    return level_offset(p, p.levels)


def level_queries(p: GiParams, level: int) -> np.ndarray:
    """Surrogate queries (position, direction, normal) of every (probe, bin)
    of a level, in merged-array order."""
    lib = _tsig()
    n = level_offset(p, level + 1) - level_offset(p, level)
    out = np.zeros((n, 9), np.float32)
    lib.orc_gi_level_queries(C.byref(p), level, _p(out))
    return out


def sample_index(seed: int, i, count: int) -> np.ndarray:
    lib = _tsig()
    return np.array([lib.orc_gi_sample_index(seed, int(k), count) for k in np.atleast_1d(i)], np.int64)


def bin_direction(index, nphi, ny, ox=0.5, oy=0.5):
    out = np.zeros(3, np.float32)
    _tsig().orc_bin_direction(index, nphi, ny, ox, oy, out.ctypes.data)
    return out


def cos_sin_2pi(u):
    c, s = C.c_float(), C.c_float()
    _tsig().orc_cos_sin_2pi(float(u), C.byref(c), C.byref(s))
    return c.value, s.value


def build_lights(scene, cap: int = 1 << 20) -> np.ndarray:
    """(L, 24) float32 light records (uint words in the last 4 columns)."""
    lib = _tsig()
    nm = len(scene.meshes)
    pos = [np.ascontiguousarray(m, np.float32) for m in scene.meshes]
    mat = [np.ascontiguousarray(m, np.uint32) for m in scene.mesh_materials]
    counts = np.asarray([p.shape[0] for p in pos], np.uint32)
    pp = (C.c_void_p * nm)(*[p.ctypes.data for p in pos])
    mp = (C.c_void_p * nm)(*[m.ctypes.data for m in mat])
    fwd = np.ascontiguousarray(scene.inst_fwd, np.float32)
    im = np.ascontiguousarray(scene.inst_mesh, np.uint32)
    fl = np.ascontiguousarray(scene.inst_flags, np.uint32)
    mats = np.ascontiguousarray(scene.materials, np.float32)
    cap = min(cap, int(counts.sum()) * max(1, im.shape[0]) + 1)
    out = np.zeros((cap, 24), np.float32)
    n = lib.orc_build_lights(pp, mp, counts.ctypes.data, nm, fwd.ctypes.data, im.ctypes.data,
                             fl.ctypes.data, im.shape[0], mats.ctypes.data, mats.shape[0],
                             out.ctypes.data, cap)
    return out[:n].copy()


def gi_trace(built: BuiltScene, lights: np.ndarray, materials: np.ndarray, p: GiParams,
             threads: int = 8, f64: bool = False, samples: bool = False, dep: bool = False):
    """Intervals (n, 4); with samples=True also the supported rays per
    interval; with dep=True also (samples and) each interval's dependency
    (n, 8: lo.xyz, known, hi.xyz, has bounds)."""
    lib = _tsig()
    lights = np.ascontiguousarray(lights, np.float32)
    mats = np.ascontiguousarray(materials, np.float32)
    out = np.zeros((total_intervals(p), 4), np.float64 if f64 else np.float32)
    if dep:
        assert not f64
        ns = np.zeros(total_intervals(p), np.uint32)
        dp = np.zeros((total_intervals(p), 8), np.float32)
        lib.orc_gi_trace_dep(C.byref(built._s), lights.ctypes.data, lights.shape[0], mats.ctypes.data,
                             mats.shape[0], C.byref(p), out.ctypes.data, ns.ctypes.data, dp.ctypes.data, threads)
        return out, ns, dp
    if samples:
        assert not f64
        ns = np.zeros(total_intervals(p), np.uint32)
        lib.orc_gi_trace_samples(C.byref(built._s), lights.ctypes.data, lights.shape[0], mats.ctypes.data,
                                 mats.shape[0], C.byref(p), out.ctypes.data, ns.ctypes.data, threads)
        return out, ns
    fn = lib.orc_gi_trace_f64 if f64 else lib.orc_gi_trace
    fn(C.byref(built._s), lights.ctypes.data, lights.shape[0], mats.ctypes.data, mats.shape[0],
       C.byref(p), out.ctypes.data, threads)
    return out


def gi_trace_stride(built: BuiltScene, lights: np.ndarray, materials: np.ndarray, p: GiParams, stride: int,
                    threads: int = 8) -> np.ndarray:
    """Intervals (n, 4) with only every stride-th one traced (the rest zero)."""
    lib = _tsig()
    lights = np.ascontiguousarray(lights, np.float32)
    mats = np.ascontiguousarray(materials, np.float32)
    out = np.zeros((total_intervals(p), 4), np.float32)
    lib.orc_gi_trace_stride(C.byref(built._s), lights.ctypes.data, lights.shape[0], mats.ctypes.data,
                            mats.shape[0], C.byref(p), out.ctypes.data, threads, stride)
    return out


def gi_merge(p: GiParams, intervals: np.ndarray, history: np.ndarray | None = None) -> np.ndarray:
    """Merge of all levels; with a history (n, 8) the masses come from it."""
    lib = _tsig()
    f64 = intervals.dtype == np.float64
    iv = np.ascontiguousarray(intervals)
    out = np.zeros_like(iv)
    if history is not None:
        h = np.ascontiguousarray(history, np.float32)
        lib.orc_gi_merge_history(C.byref(p), iv.ctypes.data, h.ctypes.data, out.ctypes.data)
        return out
    (lib.orc_gi_merge_f64 if f64 else lib.orc_gi_merge)(C.byref(p), iv.ctypes.data, out.ctypes.data)
    return out


class HistoryParams(C.Structure):
    _fields_ = [("radiance_blend", C.c_float), ("transmittance_blend", C.c_float),
                ("history_frames", C.c_uint32), ("persistent", C.c_uint32), ("valid", C.c_uint32),
                ("root_grew", C.c_uint32)]


def gi_history(p: GiParams, radiance_blend: float, transmittance_blend: float, history_frames: int,
               persistent: int, valid: bool, intervals: np.ndarray, samples: np.ndarray, history: np.ndarray,
               dep: np.ndarray, boxes=(), count: int | None = None, root_grew: bool = False):
    """One temporal history pass over intervals [0, count): returns the
    blended intervals and the new history (n, 16: J.rgb, radiance mass, beta,
    transmittance mass, last frame, 0, then the retained dependency lo.xyz,
    known, hi.xyz, has bounds); dep (n, 8) = this trace's dependency
    (gi_trace(..., dep=True)); boxes = dirty (lo, hi) pairs; root_grew: the
    scene's root box grew since the history was written (measured-empty
    dependencies are then not reused)."""
    lib = _tsig()
    h = HistoryParams(radiance_blend, transmittance_blend, history_frames, persistent, int(bool(valid)),
                      int(bool(root_grew)))
    iv = np.array(intervals, np.float32, copy=True, order="C")
    hs = np.array(history, np.float32, copy=True, order="C")
    ns = np.ascontiguousarray(samples, np.uint32)
    dp = np.ascontiguousarray(dep, np.float32)
    bx = np.zeros((max(1, len(boxes)), 6), np.float32)
    for i, (lo, hi) in enumerate(boxes):
        bx[i, :3], bx[i, 3:] = lo, hi
    n = iv.shape[0] if count is None else count
    lib.orc_gi_history(C.byref(p), C.byref(h), bx.ctypes.data, len(boxes), n, iv.ctypes.data,
                       ns.ctypes.data, dp.ctypes.data, hs.ctypes.data)
    return iv, hs


def gi_pack(p: GiParams, merged: np.ndarray) -> np.ndarray:
    lib = _tsig()
    f64 = merged.dtype == np.float64
    n0, _ = level_dims(p, 0)
    out = np.zeros((n0, 36, 4), merged.dtype)
    m = np.ascontiguousarray(merged)
    (lib.orc_gi_pack_f64 if f64 else lib.orc_gi_pack)(C.byref(p), m.ctypes.data, out.ctypes.data)
    return out


class LodView(C.Structure):
    """Mirror of orc_lod_view / lux_lod_view: camera terms of the LOD choice."""
    _fields_ = [("camera_position", C.c_float * 3), ("probe_spacing", C.c_float),
                ("clip_row1", C.c_float * 3), ("viewport_height", C.c_uint32),
                ("clip_row3", C.c_float * 3), ("target_pixels", C.c_uint32)]


def lod_view(camera=(0.0, 0.0, 0.0), probe_spacing=1.0, clip_row1=(0.0, 1.0, 0.0), viewport_height=2,
             clip_row3=(0.0, 0.0, -1.0), target_pixels=1) -> LodView:
    v = LodView()
    v.camera_position[:] = [float(x) for x in camera]
    v.probe_spacing = probe_spacing
    v.clip_row1[:] = [float(x) for x in clip_row1]
    v.viewport_height = viewport_height
    v.clip_row3[:] = [float(x) for x in clip_row3]
    v.target_pixels = target_pixels
    return v


def lod_select(view: LodView, position) -> tuple:
    lo, hi, b = C.c_uint32(), C.c_uint32(), C.c_float()
    pos = np.ascontiguousarray(position, np.float32)
    _tsig().orc_lod_select(C.byref(view), pos.ctypes.data, C.byref(lo), C.byref(hi), C.byref(b))
    return int(lo.value), int(hi.value), float(b.value)


def world_resolve_lods(grids, probes, view: LodView, receivers: np.ndarray) -> np.ndarray:
    """resolveSplitDiffuse over dense grids: grids[k] (GiParams) with packed
    probes[k] at LOD k."""
    lib = _tsig()
    f64 = probes[0].dtype == np.float64
    arrs = [np.ascontiguousarray(a) for a in probes]
    gp = (C.c_void_p * len(grids))(*[C.cast(C.byref(g), C.c_void_p) for g in grids])
    pp = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    rc = np.ascontiguousarray(receivers, np.float32)
    out = np.zeros((rc.shape[0], 4), arrs[0].dtype)
    (lib.orc_world_resolve_lods_f64 if f64 else lib.orc_world_resolve_lods)(
        gp, pp, len(grids), C.byref(view), rc.ctypes.data, rc.shape[0], out.ctypes.data)
    return out


def world_resolve(p: GiParams, probes: np.ndarray, receivers: np.ndarray) -> np.ndarray:
    lib = _tsig()
    f64 = probes.dtype == np.float64
    pr = np.ascontiguousarray(probes)
    rc = np.ascontiguousarray(receivers, np.float32)
    out = np.zeros((rc.shape[0], 4), probes.dtype)
    (lib.orc_world_resolve_f64 if f64 else lib.orc_world_resolve)(
        C.byref(p), pr.ctypes.data, rc.ctypes.data, rc.shape[0], out.ctypes.data)
    return out


# ---------------- receiver-driven probe selection ----------------
EMPTY = 0xFFFFFFFF
RX_VALID, RX_RAY = 1, 2


class ReceiverParams(C.Structure):
    """Mirror of orc_receiver_params / lux_receiver_params."""
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("budget", C.c_uint32),
                ("fair_cursor", C.c_uint32), ("fair_quota", C.c_uint32), ("frame", C.c_uint32),
                ("flags", C.c_uint32), ("reserved", C.c_uint32)]


def receiver_params(width, height, budget=0, fair_cursor=0, fair_quota=None, frame=0, disabled=False):
    rp = ReceiverParams()
    rp.width, rp.height, rp.budget = width, height, budget
    rp.fair_cursor = fair_cursor
    rp.fair_quota = min(budget, width * height) if fair_quota is None else fair_quota
    rp.frame = frame
    rp.flags = 1 if disabled else 0
    return rp


def candidate_rank(sample: int, width: int, height: int, count: int) -> int:
    return int(_tsig().orc_candidate_rank(sample, width, height, count))


def candidate_from_rank(rank: int, width: int, height: int, count: int) -> int:
    return int(_tsig().orc_candidate_from_rank(rank, width, height, count))


def candidate_selected(sample: int, rp: ReceiverParams) -> bool:
    return bool(_tsig().orc_candidate_selected(sample, C.byref(rp)))


def rx_cell(pos, base_spacing: float, level: int, lod: int) -> np.ndarray:
    k = np.zeros(4, np.int32)
    p = np.ascontiguousarray(pos, np.float32)
    _tsig().orc_rx_cell(p.ctypes.data, base_spacing, level, lod, k.ctypes.data)
    return k


def direction_index(d, nphi: int, ny: int) -> int:
    v = np.ascontiguousarray(d, np.float32)
    return int(_tsig().orc_direction_index(v.ctypes.data, nphi, ny))


def cold_candidate_rank(sample: int, width: int, height: int, count: int, frame: int) -> int:
    return int(_tsig().orc_cold_candidate_rank(sample, width, height, count, frame))


def cold_candidate_from_rank(rank: int, width: int, height: int, count: int, frame: int) -> int:
    return int(_tsig().orc_cold_candidate_from_rank(rank, width, height, count, frame))


def rx_seed(receivers: np.ndarray, rp: ReceiverParams, view, base_spacing: float, levels: int, cold: bool = False):
    """Descriptors of every (variant, receiver, level): keys (2*count*levels, 4)
    int32, flags uint32, and the LOD of each (variant, receiver) (2, count);
    cold: every receiver's selected LOD is eligible (cold priority)."""
    count = rp.width * rp.height
    rc = np.ascontiguousarray(receivers, np.float32)
    assert rc.shape == (count, 6)
    nd = 2 * count * levels
    keys = np.zeros((nd, 4), np.int32)
    flags = np.zeros(nd, np.uint32)
    lods = np.zeros((2, count), np.uint32)
    _tsig().orc_rx_seed(rc.ctypes.data, C.byref(rp), None if view is None else C.byref(view), base_spacing,
                        levels, int(cold), keys.ctypes.data, flags.ctypes.data, lods.ctypes.data)
    return keys, flags, lods


def rx_select(count: int, levels: int, keys: np.ndarray, flags: np.ndarray, rp: ReceiverParams | None = None,
              cold: bool = False) -> dict:
    """rank / parents / prefix / generateDirections; cold runs rank's
    cold-priority phases with the receiver parameters rp.  flags: the
    descriptor flags after rank (cold: level-0 ray = traced, 4 = admitted
    nominee)."""
    keys = np.ascontiguousarray(keys, np.int32)
    flags = np.ascontiguousarray(flags, np.uint32)
    nd = 2 * count * levels
    assert keys.shape == (nd, 4) and flags.shape == (nd,)
    canonical = np.zeros(nd, np.uint32)
    pcount = np.zeros(nd, np.uint32)
    poffset = np.zeros(nd, np.uint32)
    seq = np.zeros((2, count), np.uint32)
    tickets = np.zeros(max(count, 1), np.uint32)
    nt = C.c_uint32()
    fout = np.zeros(nd, np.uint32)
    assert not cold or rp is not None
    root = _tsig().orc_rx_select(count, levels, keys.ctypes.data, flags.ctypes.data,
                                 None if rp is None else C.byref(rp), int(cold), canonical.ctypes.data,
                                 pcount.ctypes.data, poffset.ctypes.data, seq.ctypes.data, tickets.ctypes.data,
                                 C.byref(nt), fout.ctypes.data)
    return dict(canonical=canonical, count=pcount, offset=poffset, sequence=seq, tickets=tickets[:nt.value],
                root=int(root), flags=fout)


def rx_trace(built: BuiltScene, lights: np.ndarray, materials: np.ndarray, grids, receivers: np.ndarray,
             lods: np.ndarray, sequence: np.ndarray, tickets: np.ndarray, frame: int):
    """Receiver trace + deposit + reduce into the grids of a LOD pyramid:
    (intervals per LOD, samples per LOD, stats[rays, unsupported, kept, outside])."""
    lib = _tsig()
    lights = np.ascontiguousarray(lights, np.float32)
    mats = np.ascontiguousarray(materials, np.float32)
    rc = np.ascontiguousarray(receivers, np.float32)
    count = rc.shape[0]
    lods = np.ascontiguousarray(lods, np.uint32)
    seq = np.ascontiguousarray(sequence, np.uint32)
    tk = np.ascontiguousarray(tickets, np.uint32)
    ivs = [np.zeros((total_intervals(g), 4), np.float32) for g in grids]
    sms = [np.zeros(total_intervals(g), np.uint32) for g in grids]
    gp = (C.c_void_p * len(grids))(*[C.cast(C.byref(g), C.c_void_p) for g in grids])
    ip = (C.c_void_p * len(grids))(*[a.ctypes.data for a in ivs])
    sp = (C.c_void_p * len(grids))(*[a.ctypes.data for a in sms])
    st = np.zeros(4, np.uint32)
    lib.orc_rx_trace(C.byref(built._s), lights.ctypes.data, lights.shape[0], mats.ctypes.data, mats.shape[0], gp,
                     len(grids), rc.ctypes.data, count, lods.ctypes.data, seq.ctypes.data, tk.ctypes.data,
                     tk.shape[0], frame, ip, sp, st.ctypes.data)
    return ivs, sms, st


# ---------------- screen-space temporal and denoise (orc_screen.c) ----------------
class ScreenParams(C.Structure):
    """orc_screen_params (oracle/lux_oracle.h) = lux_screen_params (include/lux.h)."""
    _fields_ = [
        ("world_from_clip", C.c_float * 16),
        ("previous_clip_from_world", C.c_float * 16),
        ("previous_world_from_clip", C.c_float * 16),
        ("viewport", C.c_uint32 * 4),
        ("target", C.c_uint32 * 2),
        ("size", C.c_uint32 * 2),
        ("resolve_scale", C.c_uint32),
        ("persistent", C.c_uint32),
        ("probe_spacing", C.c_float),
        ("normal_bias", C.c_float),
        ("history_blend", C.c_float),
        ("history_frames", C.c_float),
        ("stability", C.c_float * 4),
        ("source_changes", C.c_uint32),
        ("source_flags", C.c_uint32),
        ("motion_epoch", C.c_uint32),
        ("frame", C.c_uint32),
        ("dirty_min", C.c_float * 4),
        ("dirty_max", C.c_float * 4),
    ]


def _screen_sig():
    lib = load()
    vp, u32 = C.c_void_p, C.c_uint32
    lib.orc_screen_extent.argtypes = [vp, vp, vp]
    lib.orc_screen_temporal.argtypes = [vp] + [vp] * 14
    lib.orc_screen_denoise.argtypes = [vp, u32, vp, vp, vp, vp, vp, vp]
    lib.orc_screen_packed_normal.restype = C.c_float
    lib.orc_screen_packed_normal.argtypes = [vp]
    return lib


def screen_extent(P: ScreenParams):
    w, h = C.c_uint32(), C.c_uint32()
    _screen_sig().orc_screen_extent(C.byref(P), C.byref(w), C.byref(h))
    return int(w.value), int(h.value)


def screen_packed_normal(n) -> float:
    v = np.ascontiguousarray(n, np.float32)
    return float(_screen_sig().orc_screen_packed_normal(v.ctypes.data))


def _f(a, shape, dtype=np.float32):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype)
    assert a.size == int(np.prod(shape)), (a.shape, shape)
    return a


def screen_temporal(P: ScreenParams, current, current_background, depth, gbuffer, raster, motion, previous,
                    previous_metadata, previous_receiver, previous_background):
    """temporal over the resolve extent: (irradiance, metadata, receiver,
    background), each (RH, RW, 4) float32.  Full-resolution inputs are
    (target h, target w[, c]); background inputs and motion may be None."""
    RW, RH = screen_extent(P)
    TW, TH = P.target[0], P.target[1]
    r = (RH, RW, 4)
    ins = [_f(current, r), _f(current_background, r), _f(depth, (TH, TW)), _f(gbuffer, (TH, TW, 4), np.uint32),
           _f(raster, (TH, TW, 4), np.uint32), _f(motion, (TH, TW, 2)), _f(previous, r), _f(previous_metadata, r),
           _f(previous_receiver, r), _f(previous_background, r)]
    outs = [np.zeros(r, np.float32) for _ in range(4)]
    _screen_sig().orc_screen_temporal(C.byref(P), *[None if a is None else a.ctypes.data for a in ins],
                                      *[o.ctypes.data for o in outs])
    return tuple(outs)


def screen_denoise(P: ScreenParams, pass_index: int, irradiance, moments, receiver, depth):
    """One denoise pass: (irradiance, receiver), each (RH, RW, 4)."""
    RW, RH = screen_extent(P)
    r = (RH, RW, 4)
    ins = [_f(irradiance, r), _f(moments, r), _f(receiver, r), _f(depth, (P.target[1], P.target[0]))]
    out, out_r = np.zeros(r, np.float32), np.zeros(r, np.float32)
    _screen_sig().orc_screen_denoise(C.byref(P), pass_index, *[a.ctypes.data for a in ins], out.ctypes.data,
                                     out_r.ctypes.data)
    return out, out_r
