"""ctypes binding of liblux_hip.so (the C ABI declared in include/lux.h).

The library must be built (``python -m bevy_penumbra_amd.build``); loading
fails loudly when it is missing -- there is no CPU fallback for any entry
point.  Device buffers are plain integers (addresses); torch tensors are
convenient owners of device memory and are passed as ``t.data_ptr()``.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liblux_hip.so"

LUX_OK = 0
LUX_HIT_BLAS_OVERFLOW = 1
LUX_HIT_TLAS_OVERFLOW = 2
LUX_HIT_UNSUPPORTED = 4
LUX_HIT_OCCLUDED = 0x100
LUX_HIT_ROOT = 0x200
LUX_RAY_ANY_HIT = 1
LUX_RAY_FIRST_HIT = 2
LUX_EMPTY = 0xFFFFFFFF
LUX_BVH_MORTON30 = 0
LUX_BVH_FORCE_BLAS = 0x100
LUX_BVH_REFIT = 0x200
LUX_BVH_SAH = 0x1
LUX_BVH_SAH_HOST = 0x2


class LuxError(RuntimeError):
    pass


class MeshDesc(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("materials", C.c_void_p), ("triangle_count", C.c_uint32)]


class GiParams(C.Structure):
    """lux_gi_params (dense probe grid, cascades, intervals, NEE sampling)."""
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


class RadianceCascades(C.Structure):
    """C view of the Bevy RadianceCascades component (include/lux_bevy.h)."""
    _fields_ = [("method", C.c_uint32), ("history_epoch", C.c_uint32),
                ("probe_spacing", C.c_float), ("interval_length", C.c_float),
                ("resolution_divisor", C.c_uint32), ("max_rays", C.c_uint32),
                ("normal_bias", C.c_float), ("intensity", C.c_float),
                ("environment", C.c_float * 3), ("temporal_blend", C.c_float),
                ("split_transmittance_blend", C.c_float), ("history_frames", C.c_uint32),
                ("spatial_filter_steps", C.c_uint32), ("emissive_samples", C.c_uint32),
                ("sampling", C.c_uint32 * 6), ("flags", C.c_uint32)]

    @classmethod
    def default(cls) -> "RadianceCascades":
        r = cls()
        load_library().lux_radiance_cascades_default(C.byref(r))
        return r

    @classmethod
    def low_end(cls) -> "RadianceCascades":
        r = cls()
        load_library().lux_radiance_cascades_low_end(C.byref(r))
        return r

    def sanitized(self) -> "RadianceCascades":
        r = RadianceCascades()
        load_library().lux_radiance_cascades_sanitized(C.byref(self), C.byref(r))
        return r

    def gi_params(self, origin, dims, cascades, base_phi=8, base_y=4, interval_mode=0,
                  frame=0) -> "GiParams":
        p = GiParams()
        o = (C.c_float * 3)(*origin)
        d = (C.c_uint32 * 3)(*dims)
        rc = load_library().lux_radiance_cascades_gi_params(
            C.byref(self), o, d, cascades, base_phi, base_y, interval_mode, frame, C.byref(p))
        if rc != LUX_OK:
            raise LuxError(f"lux_radiance_cascades_gi_params: {rc}")
        return p

    def receiver_grid(self, viewport_width: int, viewport_height: int):
        w, h = C.c_uint32(), C.c_uint32()
        load_library().lux_radiance_cascades_receiver_grid(C.byref(self), viewport_width, viewport_height,
                                                           C.byref(w), C.byref(h))
        return int(w.value), int(h.value)

    def update_budget(self, runtime_cap=None) -> int:
        b = C.c_uint32()
        load_library().lux_radiance_cascades_update_budget(C.byref(self), runtime_cap is not None,
                                                           runtime_cap or 0, C.byref(b))
        return int(b.value)


class CandidateSchedule(C.Structure):
    """lux_candidate_schedule: the fair window's cursor (CandidateSchedule)."""
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("count", C.c_uint32),
                ("cursor", C.c_uint32), ("pending", C.c_uint32)]

    def phase_for(self, width: int, height: int, count: int, budget: int):
        c, q = C.c_uint32(), C.c_uint32()
        load_library().lux_candidate_schedule_phase(C.byref(self), width, height, count, budget, C.byref(c),
                                                    C.byref(q))
        return int(c.value), int(q.value)

    def advance_after_successful_update(self, consume_fair_cursor: bool = True):
        load_library().lux_candidate_schedule_advance(C.byref(self), int(consume_fair_cursor))


def cold_priority_enabled(history_frames: int, history_was_valid: bool) -> bool:
    """cold_priority_enabled (src/render.rs:128-131)."""
    return bool(load_library().lux_cold_priority_enabled(history_frames, int(history_was_valid)))


class SurrogateParams(C.Structure):
    """lux_surrogate_params (include/lux.h)."""
    _fields_ = [
        ("levels", C.c_uint32),
        ("features", C.c_uint32),
        ("log2_table", C.c_uint32),
        ("base_resolution", C.c_float),
        ("growth", C.c_float),
        ("box_min", C.c_float * 3),
        ("box_max", C.c_float * 3),
        ("learning_rate", C.c_float),
        ("error_weight", C.c_float),
        ("seed", C.c_uint32),
        ("freeze_updates", C.c_uint32),
        ("ema_decay", C.c_float),
        ("table_learning_rate", C.c_float),
        ("reserved", C.c_uint32 * 3),
    ]

    @classmethod
    def default(cls, box_min=(0.0, 0.0, 0.0), box_max=(1.0, 1.0, 1.0), **kw) -> "SurrogateParams":
        p = cls()
        load_library().lux_surrogate_default_params(C.byref(p))
        p.box_min[:] = [float(v) for v in box_min]
        p.box_max[:] = [float(v) for v in box_max]
        for k, v in kw.items():
            setattr(p, k, v)
        return p


class SurrogateStats(C.Structure):
    """lux_gi_surrogate_stats (include/lux.h)."""
    _fields_ = [
        ("level", C.c_uint32),
        ("probes", C.c_uint32),
        ("fallback_probes", C.c_uint32),
        ("queries", C.c_uint32),
        ("ms_total", C.c_float),
        ("ms_trace", C.c_float),
        ("ms_infer", C.c_float),
        ("ms_fallback", C.c_float),
        ("ms_merge_pack", C.c_float),
        ("reserved", C.c_float * 3),
    ]

    def as_dict(self) -> dict:
        return {k: (getattr(self, k) if not isinstance(getattr(self, k), float) else round(getattr(self, k), 4))
                for k, _ in self._fields_ if k != "reserved"}


class HistoryParams(C.Structure):
    """lux_gi_history_params (include/lux.h)."""
    _fields_ = [("radiance_blend", C.c_float), ("transmittance_blend", C.c_float),
                ("history_frames", C.c_uint32), ("persistent", C.c_uint32)]


class LodView(C.Structure):
    """lux_lod_view (include/lux.h): camera terms of the split LOD choice."""
    _fields_ = [("camera_position", C.c_float * 3), ("probe_spacing", C.c_float),
                ("clip_row1", C.c_float * 3), ("viewport_height", C.c_uint32),
                ("clip_row3", C.c_float * 3), ("target_pixels", C.c_uint32)]

    @classmethod
    def make(cls, camera=(0.0, 0.0, 0.0), probe_spacing=1.0, clip_row1=(0.0, 1.0, 0.0), viewport_height=2,
             clip_row3=(0.0, 0.0, -1.0), target_pixels=1) -> "LodView":
        v = cls()
        v.camera_position[:] = [float(x) for x in camera]
        v.probe_spacing = probe_spacing
        v.clip_row1[:] = [float(x) for x in clip_row1]
        v.viewport_height = viewport_height
        v.clip_row3[:] = [float(x) for x in clip_row3]
        v.target_pixels = target_pixels
        return v


def lod_gi_params(p: "GiParams", lod: int) -> "GiParams":
    """The grid of LOD `lod` of a LOD-0 grid: spacing and interval length
    times 2^lod (shaders/common.slang:414-416, shaders/split.slang:224-227),
    dims halved (rounded up) per LOD, same origin."""
    q = GiParams()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(p))
    q.spacing = p.spacing * 2.0 ** lod
    q.interval_length = p.interval_length * 2.0 ** lod
    q.dims[:] = [max(1, -(-int(d) // (1 << lod))) for d in p.dims]
    return q


def world_resolve_lods(contexts, view: LodView, receivers, n: int, out):
    """lux_world_resolve_lods: resolve against the LOD pyramid contexts[k] = LOD k."""
    lib = contexts[0].lib
    arr = (C.c_void_p * len(contexts))(*[c.h for c in contexts])
    contexts[0]._check(lib.lux_world_resolve_lods(arr, len(contexts), C.byref(view), _ptr(receivers), n, _ptr(out)),
                       "lux_world_resolve_lods")


class ReceiverParams(C.Structure):
    """lux_receiver_params (include/lux.h)."""
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("budget", C.c_uint32),
                ("fair_cursor", C.c_uint32), ("fair_quota", C.c_uint32), ("frame", C.c_uint32),
                ("flags", C.c_uint32), ("reserved", C.c_uint32)]


class ReceiverStats(C.Structure):
    """lux_receiver_stats (include/lux.h)."""
    _fields_ = [("receivers", C.c_uint32), ("tickets", C.c_uint32), ("rays", C.c_uint32),
                ("unsupported", C.c_uint32), ("deposits", C.c_uint32), ("outside", C.c_uint32),
                ("root_rays", C.c_uint32), ("overflow", C.c_uint32), ("ms", C.c_float),
                ("cold", C.c_uint32), ("nominees", C.c_uint32), ("reserved", C.c_uint32)]


def receiver_params(width: int, height: int, budget: int = 0, fair_cursor: int = 0, fair_quota=None,
                    frame: int = 0, disabled: bool = False) -> ReceiverParams:
    rp = ReceiverParams()
    rp.width, rp.height, rp.budget = width, height, budget
    rp.fair_cursor = fair_cursor
    rp.fair_quota = min(budget, width * height) if fair_quota is None else fair_quota
    rp.frame = frame
    rp.flags = 1 if disabled else 0
    return rp


class ScreenParams(C.Structure):
    """lux_screen_params (include/lux.h): the temporal / denoise parameters."""
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

    def extent(self):
        w, h = C.c_uint32(), C.c_uint32()
        load_library().lux_screen_extent(C.byref(self), C.byref(w), C.byref(h))
        return int(w.value), int(h.value)


def screen_temporal(ctx, params: ScreenParams, current, current_background, depth, gbuffer, raster, motion,
                    previous, previous_metadata, previous_receiver, previous_background, irradiance, metadata,
                    receiver, background):
    """lux_screen_temporal on device buffers (None for the nullable ones)."""
    ctx._check(ctx.lib.lux_screen_temporal(ctx.h, C.byref(params), *[_ptr(b) for b in (
        current, current_background, depth, gbuffer, raster, motion, previous, previous_metadata,
        previous_receiver, previous_background, irradiance, metadata, receiver, background)]),
        "lux_screen_temporal")


def screen_denoise(ctx, params: ScreenParams, pass_index: int, irradiance, moments, receiver, depth, out,
                   out_receiver):
    """lux_screen_denoise: one pass on device buffers."""
    ctx._check(ctx.lib.lux_screen_denoise(ctx.h, C.byref(params), pass_index, *[_ptr(b) for b in (
        irradiance, moments, receiver, depth, out, out_receiver)]), "lux_screen_denoise")


def update_receivers(contexts, view, receivers, params: ReceiverParams) -> ReceiverStats:
    """lux_gi_update_receivers over the LOD pyramid contexts[k] = LOD k; view
    None: LOD 0 for every receiver."""
    lib = contexts[0].lib
    arr = (C.c_void_p * len(contexts))(*[c.h for c in contexts])
    st = ReceiverStats()
    contexts[0]._check(lib.lux_gi_update_receivers(arr, len(contexts), None if view is None else C.byref(view),
                                                   _ptr(receivers), C.byref(params), C.byref(st)),
                       "lux_gi_update_receivers")
    return st


def receivers_download(ctx, count: int, levels: int, directory: bool = True) -> dict:
    """The last receiver update's selection on contexts[0] (see
    lux_receivers_download); directory=False skips the per-descriptor keys,
    canonical indices, counts and offsets."""
    nd = 2 * count * levels
    keys = np.zeros((nd, 4), np.int32) if directory else None
    flags = np.zeros(nd, np.uint32)
    canon = np.zeros(nd, np.uint32) if directory else None
    cnt = np.zeros(nd, np.uint32) if directory else None
    off = np.zeros(nd, np.uint32) if directory else None
    lods = np.zeros((2, count), np.uint32)
    seq = np.zeros((2, count), np.uint32)
    tickets = np.zeros(max(count, 1), np.uint32)
    nt = C.c_uint32()
    ctx._check(ctx.lib.lux_receivers_download(ctx.h, _ptr(keys), _ptr(flags), _ptr(canon), _ptr(cnt), _ptr(off),
                                              _ptr(lods), _ptr(seq), _ptr(tickets), C.byref(nt)),
               "lux_receivers_download")
    return dict(keys=keys, flags=flags, canonical=canon, count=cnt, offset=off, lods=lods, sequence=seq,
                tickets=tickets[:nt.value])


LUX_SAMPLES_TRACE = 0
LUX_SAMPLES_MERGED = 1


class InstanceDesc(C.Structure):
    _fields_ = [
        ("local_from_world", C.c_float * 12),
        ("world_from_local", C.c_float * 12),
        ("mesh", C.c_uint32),
        ("flags", C.c_uint32),
    ]


_lib = None


def load_library() -> C.CDLL:
    """Load liblux_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise LuxError(
            f"{LIB_PATH} is missing: build it with `python -m bevy_penumbra_amd.build` "
            "(the HIP path has no fallback)"
        )
    lib = C.CDLL(str(LIB_PATH))
    vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int
    sig = {
        "lux_ctx_create": (i32, [i32, C.POINTER(vp)]),
        "lux_ctx_destroy": (None, [vp]),
        "lux_last_error": (C.c_char_p, [vp]),
        "lux_ctx_stream": (vp, [vp]),
        "lux_ctx_synchronize": (i32, [vp]),
        "lux_scene_upload_soa": (i32, [vp, vp, u32, vp, u32, vp, u32]),
        "lux_scene_set_transforms": (i32, [vp, vp, vp, u32]),
        "lux_scene_upload_built": (i32, [vp, vp, u32, vp, u32, vp, u32, vp, u32]),
        "lux_bvh_build": (i32, [vp, u32]),
        "lux_bvh_info": (i32, [vp, vp, vp, vp, vp]),
        "lux_bvh_download": (i32, [vp, vp, vp, vp, vp]),
        "lux_sort_pairs_device": (i32, [vp, vp, u32, vp, vp]),
        "lux_morton_codes_device": (i32, [vp, vp, vp, u32, vp]),
        "lux_trace_rays": (i32, [vp, vp, u32, vp]),
        "lux_trace_rays_device": (i32, [vp, vp, u32, vp, vp]),
        "lux_trace_rays_device_noalloc": (i32, [vp, vp, u32, vp, vp]),
        "lux_check_reciprocal_division": (i32, [vp, C.c_uint64, u32, C.POINTER(C.c_uint64)]),
        "lux_trace_stats_device": (i32, [vp, vp, u32, vp, vp]),
        "lux_last_kernel_ms": (i32, [vp, C.POINTER(C.c_float)]),
        "lux_set_trace_tuning": (i32, [vp, u32, u32]),
        "lux_gi_configure": (i32, [vp, vp]),
        "lux_gi_set_stage_timing": (i32, [vp, i32]),
        "lux_gi_stage_times": (i32, [vp, vp]),
        "lux_gi_trace_stats": (i32, [vp, vp]),
        "lux_gi_set_frame": (i32, [vp, u32, u32]),
        "lux_gi_trace": (i32, [vp]),
        "lux_gi_merge": (i32, [vp]),
        "lux_gi_pack": (i32, [vp]),
        "lux_gi_update": (i32, [vp, vp]),
        "lux_gi_sizes": (i32, [vp, vp, vp, vp]),
        "lux_gi_download": (i32, [vp, vp, vp, vp, vp]),
        "lux_gi_upload_intervals": (i32, [vp, vp]),
        "lux_gi_device_arrays": (i32, [vp, vp, vp, vp]),
        "lux_world_resolve_device": (i32, [vp, vp, u32, vp]),
        "lux_device_alloc": (i32, [vp, C.c_size_t, C.POINTER(vp)]),
        "lux_device_free": (i32, [vp, vp]),
        "lux_copy_to_device": (i32, [vp, vp, vp, C.c_size_t]),
        "lux_copy_to_host": (i32, [vp, vp, vp, C.c_size_t]),
        "lux_device_zero": (i32, [vp, vp, C.c_size_t]),
        "lux_device_copy": (i32, [vp, vp, vp, C.c_size_t]),
        "lux_radiance_cascades_default": (i32, [vp]),
        "lux_radiance_cascades_low_end": (i32, [vp]),
        "lux_radiance_cascades_sanitized": (i32, [vp, vp]),
        "lux_radiance_cascades_gi_params": (i32, [vp, vp, vp, u32, u32, u32, u32, u32, vp]),
        "lux_gi_trace_levels": (i32, [vp, u32, u32]),
        "lux_gi_set_trace_mode": (i32, [vp, u32]),
        "lux_gi_set_trace_config": (i32, [vp, vp]),
        "lux_gi_get_trace_config": (i32, [vp, vp]),
        "lux_autotune_gi": (i32, [vp, u32, vp, vp, vp]),
        "lux_gi_query_counters": (i32, [vp, vp, vp]),
        "lux_camera_receivers": (i32, [vp, vp, u32, u32, vp]),
        "lux_set_trace_config": (i32, [vp, vp]),
        "lux_get_trace_config": (i32, [vp, vp]),
        "lux_autotune_trace": (i32, [vp, vp, u32, vp, u32, vp, vp, vp]),
        "lux_gi_reset_query_counters": (i32, [vp]),
        "lux_surrogate_default_params": (None, [vp]),
        "lux_surrogate_create": (i32, [vp, vp, C.POINTER(vp)]),
        "lux_surrogate_destroy": (None, [vp]),
        "lux_surrogate_reset": (i32, [vp]),
        "lux_surrogate_encode": (i32, [vp, vp, u32, vp]),
        "lux_surrogate_infer": (i32, [vp, vp, u32, vp]),
        "lux_surrogate_train_step": (i32, [vp, vp, u32, vp]),
        "lux_surrogate_zero_grad": (i32, [vp]),
        "lux_surrogate_sizes": (i32, [vp, vp, vp, vp]),
        "lux_surrogate_download": (i32, [vp, vp, vp, vp, vp]),
        "lux_surrogate_upload": (i32, [vp, vp, vp]),
        "lux_gi_surrogate_samples": (i32, [vp, u32, u32, u32, u32, vp]),
        "lux_gi_surrogate_queries": (i32, [vp, u32, vp, vp]),
        "lux_gi_update_surrogate": (i32, [vp, vp, u32, C.c_float, vp]),
        "lux_gi_surrogate_probe_error": (i32, [vp, vp, u32]),
        "lux_world_resolve_lods": (i32, [vp, u32, vp, vp, u32, vp]),
        "lux_scene_upload_words": (i32, [vp, vp, u32, vp, u32, vp, u32]),
        "lux_gi_set_history": (i32, [vp, vp]),
        "lux_gi_reset_history": (i32, [vp]),
        "lux_gi_invalidate_history": (i32, [vp, vp, vp]),
        "lux_gi_history_arrays": (i32, [vp, vp, vp]),
        "lux_gi_history_reused": (i32, [vp, C.POINTER(C.c_uint32)]),
        "lux_radiance_cascades_history_params": (i32, [vp, vp]),
        "lux_radiance_cascades_receiver_grid": (i32, [vp, u32, u32, vp, vp]),
        "lux_radiance_cascades_update_budget": (i32, [vp, u32, u32, vp]),
        "lux_candidate_schedule_phase": (i32, [vp, u32, u32, u32, u32, vp, vp]),
        "lux_candidate_schedule_advance": (i32, [vp, u32]),
        "lux_cold_priority_enabled": (i32, [u32, u32]),
        "lux_screen_extent": (i32, [vp, vp, vp]),
        "lux_screen_temporal": (i32, [vp, vp] + [vp] * 14),
        "lux_screen_denoise": (i32, [vp, vp, u32] + [vp] * 6),
        "lux_gi_update_receivers": (i32, [vp, u32, vp, vp, vp, vp]),
        "lux_receivers_download": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "lux_gi_samples_download": (i32, [vp, vp]),
        "lux_batch_create": (i32, [vp, u32, C.POINTER(vp)]),
        "lux_batch_destroy": (None, [vp]),
        "lux_batch_last_error": (C.c_char_p, [vp]),
        "lux_batch_set_surrogate": (i32, [vp, u32, vp, u32]),
        "lux_batch_set_surrogate_bound": (i32, [vp, u32, C.c_float]),
        "lux_batch_refresh": (i32, [vp]),
        "lux_batch_set_frame": (i32, [vp, u32, u32]),
        "lux_batch_update": (i32, [vp, vp]),
        "lux_batch_capture": (i32, [vp]),
        "lux_batch_launch": (i32, [vp, vp]),
        "lux_batch_synchronize": (i32, [vp]),
        "lux_batch_sizes": (i32, [vp, vp, vp, vp, vp]),
        "lux_batch_set_config": (i32, [vp, vp]),
        "lux_batch_get_config": (i32, [vp, vp]),
        "lux_batch_autotune": (i32, [vp, u32, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class DeviceBuffer:
    """Device memory owned through the lux ABI (no second HIP runtime)."""

    def __init__(self, ctx: "Context", nbytes: int, dtype=np.uint8, shape=None):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = shape if shape is not None else (self.nbytes // self.dtype.itemsize,)
        p = C.c_void_p()
        ctx._check(ctx.lib.lux_device_alloc(ctx.h, self.nbytes, C.byref(p)), "lux_device_alloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, ctx: "Context", a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(ctx, a.nbytes, a.dtype, a.shape)
        b.upload(a)
        return b

    @classmethod
    def empty(cls, ctx: "Context", shape, dtype) -> "DeviceBuffer":
        shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        b = cls(ctx, nbytes, dtype, shape)
        b.zero()
        return b

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        self.ctx._check(self.ctx.lib.lux_copy_to_device(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes),
                        "lux_copy_to_device")

    def copy_from(self, other: "DeviceBuffer", nbytes: int | None = None):
        n = min(self.nbytes, other.nbytes) if nbytes is None else nbytes
        self.ctx._check(self.ctx.lib.lux_device_copy(self.ctx.h, self.ptr, other.ptr, n), "lux_device_copy")

    def zero(self):
        self.ctx._check(self.ctx.lib.lux_device_zero(self.ctx.h, self.ptr, self.nbytes), "lux_device_zero")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        self.ctx._check(self.ctx.lib.lux_copy_to_host(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes),
                        "lux_copy_to_host")
        return out

    def free(self):
        if self.ptr and getattr(self.ctx, "h", None):
            self.ctx.lib.lux_device_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr(a) -> int | None:
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if isinstance(a, DeviceBuffer):
        return a.ptr
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"], "arrays passed to lux must be C-contiguous"
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    raise TypeError(type(a))


class Context:
    """One lux context (device, stream, scene, BVH)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.lux_ctx_create(device, C.byref(h))
        if rc != LUX_OK:
            raise LuxError(f"lux_ctx_create({device}) failed with {rc}")
        self.h = h
        self._keep = []

    def close(self):
        if getattr(self, "h", None):
            self.lib.lux_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int, what: str):
        if rc != LUX_OK:
            msg = self.lib.lux_last_error(self.h)
            raise LuxError(f"{what}: {rc} ({msg.decode() if msg else ''})")

    @property
    def stream(self) -> int:
        return self.lib.lux_ctx_stream(self.h) or 0

    def synchronize(self):
        self._check(self.lib.lux_ctx_synchronize(self.h), "lux_ctx_synchronize")

    # ---- scene ----
    def upload_scene(self, scene) -> None:
        meshes = (MeshDesc * max(1, len(scene.meshes)))()
        keep = []
        for i, (pos, mat) in enumerate(zip(scene.meshes, scene.mesh_materials)):
            pos = np.ascontiguousarray(pos, dtype=np.float32)
            mat = np.ascontiguousarray(mat, dtype=np.uint32)
            keep += [pos, mat]
            meshes[i].positions = pos.ctypes.data
            meshes[i].materials = mat.ctypes.data
            meshes[i].triangle_count = pos.shape[0]
        ni = scene.inst_mesh.shape[0]
        insts = (InstanceDesc * max(1, ni))()
        for i in range(ni):
            insts[i].local_from_world[:] = [float(x) for x in scene.inst_inv[i]]
            insts[i].world_from_local[:] = [float(x) for x in scene.inst_fwd[i]]
            insts[i].mesh = int(scene.inst_mesh[i])
            insts[i].flags = int(scene.inst_flags[i])
        mats = np.ascontiguousarray(scene.materials, dtype=np.float32)
        rc = self.lib.lux_scene_upload_soa(
            self.h, C.addressof(meshes), len(scene.meshes), C.addressof(insts), ni,
            mats.ctypes.data, mats.shape[0])
        self._check(rc, "lux_scene_upload_soa")

    def set_transforms(self, inv: np.ndarray, fwd: np.ndarray) -> None:
        inv = np.ascontiguousarray(inv, dtype=np.float32)
        fwd = np.ascontiguousarray(fwd, dtype=np.float32)
        self._check(self.lib.lux_scene_set_transforms(self.h, inv.ctypes.data, fwd.ctypes.data,
                                                       inv.shape[0]), "lux_scene_set_transforms")

    def upload_built(self, blas: np.ndarray, tlas: np.ndarray, tri: np.ndarray, inst: np.ndarray):
        """Install a built scene: nodes (N,8) u32 words, tri (T,12) f32, inst (I,32) u32 words."""
        arrs = [np.ascontiguousarray(a) for a in (blas, tlas, tri, inst)]
        b, t, tr, ins = arrs
        rc = self.lib.lux_scene_upload_built(self.h, b.ctypes.data, b.shape[0], t.ctypes.data,
                                             t.shape[0], tr.ctypes.data, tr.shape[0],
                                             ins.ctypes.data, ins.shape[0])
        self._check(rc, "lux_scene_upload_built")

    def upload_words(self, scene_words: np.ndarray, geometry_words: np.ndarray, materials: np.ndarray):
        """Install a scene from the reference's scene and geometry words ((n, 4) uint32 each) and an
        (M, 8) float32 material table (albedo.rgb, roughness, emission.rgb, pad)."""
        sw = np.ascontiguousarray(scene_words, np.uint32)
        gw = np.ascontiguousarray(geometry_words, np.uint32)
        mt = np.ascontiguousarray(materials, np.float32)
        self._check(self.lib.lux_scene_upload_words(self.h, sw.ctypes.data, sw.shape[0], gw.ctypes.data, gw.shape[0],
                                                    mt.ctypes.data, mt.shape[0]), "lux_scene_upload_words")

    def build_bvh(self, mode: int = LUX_BVH_MORTON30) -> None:
        self._check(self.lib.lux_bvh_build(self.h, mode), "lux_bvh_build")

    def bvh_info(self) -> dict:
        v = [C.c_uint32() for _ in range(4)]
        self._check(self.lib.lux_bvh_info(self.h, *[C.byref(x) for x in v]), "lux_bvh_info")
        return dict(blas_nodes=v[0].value, tlas_nodes=v[1].value, triangles=v[2].value,
                    instances=v[3].value)

    def download_bvh(self) -> dict:
        info = self.bvh_info()
        blas = np.zeros((info["blas_nodes"], 8), np.uint32)
        tlas = np.zeros((max(1, info["tlas_nodes"]), 8), np.uint32)
        tri_order = np.zeros(max(1, info["triangles"]), np.uint32)
        inst_order = np.zeros(max(1, info["instances"]), np.uint32)
        self._check(self.lib.lux_bvh_download(self.h, blas.ctypes.data, tlas.ctypes.data,
                                              tri_order.ctypes.data, inst_order.ctypes.data),
                    "lux_bvh_download")
        return dict(blas=blas, tlas=tlas, tri_order=tri_order[: info["triangles"]],
                    inst_order=inst_order[: info["instances"]])

    # ---- primitives ----
    def sort_pairs_device(self, keys_in, n, keys_out, index_out):
        self._check(self.lib.lux_sort_pairs_device(self.h, _ptr(keys_in), n, _ptr(keys_out),
                                                   _ptr(index_out)), "lux_sort_pairs_device")

    def morton_codes_device(self, lo4, hi4, n, codes):
        self._check(self.lib.lux_morton_codes_device(self.h, _ptr(lo4), _ptr(hi4), n, _ptr(codes)),
                    "lux_morton_codes_device")

    # ---- traversal ----
    def trace(self, rays: np.ndarray) -> np.ndarray:
        """Host rays (N,8) float32 -> hits (N,4) uint32 (t bits, tri, inst, flags)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = rays.shape[0]
        hits = np.zeros((n, 4), np.uint32)
        self._check(self.lib.lux_trace_rays(self.h, rays.ctypes.data, n, hits.ctypes.data),
                    "lux_trace_rays")
        return hits

    def trace_device(self, rays, n: int, hits, bary=None):
        self._check(self.lib.lux_trace_rays_device(self.h, _ptr(rays), n, _ptr(hits), _ptr(bary)),
                    "lux_trace_rays_device")

    def check_reciprocal_division(self, pairs: int, seed: int = 0) -> int:
        """Mismatches of the walk's reciprocal slab division against IEEE division."""
        bad = C.c_uint64()
        self._check(self.lib.lux_check_reciprocal_division(self.h, pairs, seed, C.byref(bad)),
                    "lux_check_reciprocal_division")
        return int(bad.value)

    def trace_device_noalloc(self, rays, n: int, hits, bary=None):
        """lux_trace_rays_device_noalloc: the same trace with no allocation."""
        self._check(self.lib.lux_trace_rays_device_noalloc(self.h, _ptr(rays), n, _ptr(hits), _ptr(bary)),
                    "lux_trace_rays_device_noalloc")

    def trace_stats_device(self, rays, n: int, hits, counters):
        self._check(self.lib.lux_trace_stats_device(self.h, _ptr(rays), n, _ptr(hits),
                                                    _ptr(counters)), "lux_trace_stats_device")

    def last_kernel_ms(self) -> float:
        v = C.c_float()
        self._check(self.lib.lux_last_kernel_ms(self.h, C.byref(v)), "lux_last_kernel_ms")
        return float(v.value)

    def set_trace_tuning(self, waves_per_cu: int = 0, regroup: int = 0):
        self._check(self.lib.lux_set_trace_tuning(self.h, waves_per_cu, regroup),
                    "lux_set_trace_tuning")

    def set_trace_config(self, waves_per_cu: int = 0, lds_stack_depth: int = 0,
                         regroup_threshold: int = 0, steps_per_refill: int = 0):
        cfg = (C.c_uint32 * 8)(waves_per_cu, lds_stack_depth, regroup_threshold, steps_per_refill,
                               0, 0, 0, 0)
        self._check(self.lib.lux_set_trace_config(self.h, cfg), "lux_set_trace_config")

    def autotune_trace(self, rays, n: int, hits, repeats: int = 3) -> dict:
        cfg = (C.c_uint32 * 8)()
        ms = C.c_float()
        ev = C.c_uint32()
        self._check(self.lib.lux_autotune_trace(self.h, _ptr(rays), n, _ptr(hits), repeats, cfg,
                                                C.byref(ms), C.byref(ev)), "lux_autotune_trace")
        return dict(waves_per_cu=cfg[0], lds_stack_depth=cfg[1], regroup_threshold=cfg[2],
                    steps_per_refill=cfg[3], ms=float(ms.value), evaluations=int(ev.value))

    def trace_config(self) -> dict:
        cfg = (C.c_uint32 * 8)()
        self._check(self.lib.lux_get_trace_config(self.h, cfg), "lux_get_trace_config")
        return dict(waves_per_cu=cfg[0], lds_stack_depth=cfg[1], regroup_threshold=cfg[2],
                    steps_per_refill=cfg[3])

    # ---- transport (dense-grid GI update) ----
    def gi_configure(self, params: GiParams):
        self._gi = params
        self._check(self.lib.lux_gi_configure(self.h, C.byref(params)), "lux_gi_configure")

    def gi_set_stage_timing(self, on: bool):
        self._check(self.lib.lux_gi_set_stage_timing(self.h, 1 if on else 0), "lux_gi_set_stage_timing")

    def gi_stage_times(self) -> dict:
        """Milliseconds of the last traced update's wavefront stages (stage timing on)."""
        ms = (C.c_float * 4)()
        self._check(self.lib.lux_gi_stage_times(self.h, ms), "lux_gi_stage_times")
        return dict(primary=ms[0], setup_pack=ms[1], occlusion=ms[2], accumulate=ms[3])

    def gi_trace_stats(self) -> dict:
        """Traversal work of one instrumented interval trace (lux_gi_trace_stats)."""
        c = (C.c_uint64 * 8)()
        self._check(self.lib.lux_gi_trace_stats(self.h, c), "lux_gi_trace_stats")
        keys = ("primary_walks", "primary_nodes", "primary_tris", "primary_insts",
                "occlusion_rays", "occlusion_nodes", "occlusion_tris", "occlusion_insts")
        return {k: int(v) for k, v in zip(keys, c)}

    def gi_set_frame(self, frame: int, seed: int):
        self._check(self.lib.lux_gi_set_frame(self.h, frame, seed), "lux_gi_set_frame")

    def gi_trace(self):
        self._check(self.lib.lux_gi_trace(self.h), "lux_gi_trace")

    def gi_merge(self):
        self._check(self.lib.lux_gi_merge(self.h), "lux_gi_merge")

    def gi_pack(self):
        self._check(self.lib.lux_gi_pack(self.h), "lux_gi_pack")

    def gi_update(self, timed: bool = False):
        ms = C.c_float()
        self._check(self.lib.lux_gi_update(self.h, C.byref(ms) if timed else None), "lux_gi_update")
        return float(ms.value) if timed else None

    def gi_sizes(self):
        n = C.c_uint64()
        p = C.c_uint32()
        l = C.c_uint32()
        self._check(self.lib.lux_gi_sizes(self.h, C.byref(n), C.byref(p), C.byref(l)), "lux_gi_sizes")
        return int(n.value), int(p.value), int(l.value)

    def gi_samples_download(self) -> np.ndarray:
        n, _, _ = self.gi_sizes()
        out = np.zeros(n, np.uint32)
        self._check(self.lib.lux_gi_samples_download(self.h, out.ctypes.data), "lux_gi_samples_download")
        return out

    def gi_download(self):
        n, p, l = self.gi_sizes()
        iv = np.zeros((n, 4), np.float32)
        mg = np.zeros((n, 4), np.float32)
        pr = np.zeros((p, 36, 4), np.float32)
        li = np.zeros((max(1, l), 24), np.float32)
        self._check(self.lib.lux_gi_download(self.h, iv.ctypes.data, mg.ctypes.data, pr.ctypes.data,
                                             li.ctypes.data), "lux_gi_download")
        return dict(intervals=iv, merged=mg, probes=pr, lights=li[:l])

    def gi_download_probes(self) -> np.ndarray:
        """The packed level-0 probes only, (probes, 36, 4) float32."""
        _, p, _ = self.gi_sizes()
        pr = np.zeros((p, 36, 4), np.float32)
        self._check(self.lib.lux_gi_download(self.h, None, None, pr.ctypes.data, None), "lux_gi_download")
        return pr

    def gi_upload_intervals(self, intervals: np.ndarray):
        iv = np.ascontiguousarray(intervals, np.float32)
        self._check(self.lib.lux_gi_upload_intervals(self.h, iv.ctypes.data), "lux_gi_upload_intervals")

    def gi_device_arrays(self):
        a, b, c = C.c_void_p(), C.c_void_p(), C.c_void_p()
        self._check(self.lib.lux_gi_device_arrays(self.h, C.byref(a), C.byref(b), C.byref(c)),
                    "lux_gi_device_arrays")
        return a.value, b.value, c.value

    def world_resolve_device(self, receivers, n: int, out):
        self._check(self.lib.lux_world_resolve_device(self.h, _ptr(receivers), n, _ptr(out)),
                    "lux_world_resolve_device")

    def camera_receivers(self, camera: np.ndarray, width: int, height: int, out):
        cam = np.ascontiguousarray(camera, np.float32)
        assert cam.size == 12
        self._check(self.lib.lux_camera_receivers(self.h, cam.ctypes.data, width, height, _ptr(out)),
                    "lux_camera_receivers")

    def gi_query_counters(self):
        a, b = C.c_uint32(), C.c_uint32()
        self._check(self.lib.lux_gi_query_counters(self.h, C.byref(a), C.byref(b)), "lux_gi_query_counters")
        return int(a.value), int(b.value)

    def gi_reset_query_counters(self):
        self._check(self.lib.lux_gi_reset_query_counters(self.h), "lux_gi_reset_query_counters")

    # ---- learned far-cascade surrogate ----
    def gi_set_trace_mode(self, mode: int):
        """0 wavefront (default), 1 fused; results are identical."""
        self._check(self.lib.lux_gi_set_trace_mode(self.h, mode), "lux_gi_set_trace_mode")

    def gi_set_trace_config(self, waves_per_cu: int = 24, lds_stack_depth: int = 12,
                            regroup_threshold: int = 12, steps_per_refill: int = 24):
        cfg = (C.c_uint32 * 8)(waves_per_cu, lds_stack_depth, regroup_threshold, steps_per_refill,
                               0, 0, 0, 0)
        self._check(self.lib.lux_gi_set_trace_config(self.h, cfg), "lux_gi_set_trace_config")

    def gi_trace_config(self) -> dict:
        cfg = (C.c_uint32 * 8)()
        self._check(self.lib.lux_gi_get_trace_config(self.h, cfg), "lux_gi_get_trace_config")
        return dict(waves_per_cu=cfg[0], lds_stack_depth=cfg[1], regroup_threshold=cfg[2],
                    steps_per_refill=cfg[3])

    def autotune_gi(self, repeats: int = 2) -> dict:
        cfg = (C.c_uint32 * 8)()
        ms = C.c_float()
        ev = C.c_uint32()
        self._check(self.lib.lux_autotune_gi(self.h, repeats, cfg, C.byref(ms), C.byref(ev)),
                    "lux_autotune_gi")
        return dict(waves_per_cu=cfg[0], lds_stack_depth=cfg[1], regroup_threshold=cfg[2],
                    steps_per_refill=cfg[3], ms=float(ms.value), evaluations=int(ev.value))

    def gi_trace_levels(self, first: int, last: int):
        self._check(self.lib.lux_gi_trace_levels(self.h, first, last), "lux_gi_trace_levels")

    def gi_surrogate_samples(self, level: int, n: int, sample_seed: int, source: int, out):
        self._check(self.lib.lux_gi_surrogate_samples(self.h, level, n, sample_seed, source, _ptr(out)),
                    "lux_gi_surrogate_samples")

    def gi_surrogate_query_count(self, level: int) -> int:
        c = C.c_uint64()
        self._check(self.lib.lux_gi_surrogate_queries(self.h, level, None, C.byref(c)),
                    "lux_gi_surrogate_queries")
        return int(c.value)

    def gi_surrogate_queries(self, level: int, out):
        self._check(self.lib.lux_gi_surrogate_queries(self.h, level, _ptr(out), None),
                    "lux_gi_surrogate_queries")

    def gi_update_surrogate(self, surrogate: "Surrogate", level: int, error_bound: float,
                            stats: bool = True):
        st = SurrogateStats()
        self._check(self.lib.lux_gi_update_surrogate(self.h, surrogate.h, level, float(error_bound),
                                                     C.byref(st) if stats else None),
                    "lux_gi_update_surrogate")
        return st if stats else None

    def gi_set_history(self, radiance_blend: float = 0.88, transmittance_blend: float = 0.98,
                       history_frames: int = 24, persistent: int = 1):
        hp = HistoryParams(radiance_blend, transmittance_blend, history_frames, persistent)
        self._check(self.lib.lux_gi_set_history(self.h, C.byref(hp)), "lux_gi_set_history")

    def gi_history_off(self):
        self._check(self.lib.lux_gi_set_history(self.h, None), "lux_gi_set_history")

    def gi_reset_history(self):
        self._check(self.lib.lux_gi_reset_history(self.h), "lux_gi_reset_history")

    def gi_invalidate_history(self, lo, hi):
        a = (C.c_float * 3)(*[float(x) for x in lo])
        b = (C.c_float * 3)(*[float(x) for x in hi])
        self._check(self.lib.lux_gi_invalidate_history(self.h, a, b), "lux_gi_invalidate_history")

    def gi_history_reused(self) -> int:
        """Intervals whose history the last history pass reused."""
        r = C.c_uint32()
        self._check(self.lib.lux_gi_history_reused(self.h, C.byref(r)), "lux_gi_history_reused")
        return int(r.value)

    def gi_history_download(self):
        """(history (n, 16) float32, samples (n,) uint32) of the last update:
        J.rgb, radiance mass, beta, transmittance mass, last frame, 0, then the
        retained dependency (lo.xyz, known, hi.xyz, has bounds)."""
        hp, sp = C.c_void_p(), C.c_void_p()
        self._check(self.lib.lux_gi_history_arrays(self.h, C.byref(hp), C.byref(sp)), "lux_gi_history_arrays")
        if not hp.value:
            return None, None
        n = self.gi_sizes()[0]
        self.synchronize()
        hist = np.zeros((n, 16), np.float32)
        ns = np.zeros(n, np.uint32)
        self._check(self.lib.lux_copy_to_host(self.h, hist.ctypes.data, hp.value, hist.nbytes), "lux_copy_to_host")
        self._check(self.lib.lux_copy_to_host(self.h, ns.ctypes.data, sp.value, ns.nbytes), "lux_copy_to_host")
        return hist, ns

    def gi_surrogate_probe_error(self, probes: int) -> np.ndarray:
        out = np.zeros(probes, np.float32)
        self._check(self.lib.lux_gi_surrogate_probe_error(self.h, out.ctypes.data, probes),
                    "lux_gi_surrogate_probe_error")
        return out


class Surrogate:
    """Hash-grid + MLP radiance surrogate owned by a Context (lux_surrogate)."""

    def __init__(self, ctx: Context, params: SurrogateParams):
        self.ctx = ctx
        self.lib = ctx.lib
        self.params = params
        h = C.c_void_p()
        ctx._check(self.lib.lux_surrogate_create(ctx.h, C.byref(params), C.byref(h)), "lux_surrogate_create")
        self.h = h
        m, t = C.c_uint64(), C.c_uint64()
        self.lib.lux_surrogate_sizes(self.h, C.byref(m), C.byref(t), None)
        self.mlp_params, self.table_params = int(m.value), int(t.value)

    def close(self):
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            self.lib.lux_surrogate_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def steps(self) -> int:
        s = C.c_uint32()
        self.lib.lux_surrogate_sizes(self.h, None, None, C.byref(s))
        return int(s.value)

    def reset(self):
        self.ctx._check(self.lib.lux_surrogate_reset(self.h), "lux_surrogate_reset")

    def encode(self, queries, n: int, out):
        self.ctx._check(self.lib.lux_surrogate_encode(self.h, _ptr(queries), n, _ptr(out)),
                        "lux_surrogate_encode")

    def infer(self, queries, n: int, out):
        self.ctx._check(self.lib.lux_surrogate_infer(self.h, _ptr(queries), n, _ptr(out)),
                        "lux_surrogate_infer")

    def train_step(self, samples, n: int, want_loss: bool = True):
        loss = C.c_float()
        self.ctx._check(self.lib.lux_surrogate_train_step(self.h, _ptr(samples), n,
                                                          C.byref(loss) if want_loss else None),
                        "lux_surrogate_train_step")
        return float(loss.value) if want_loss else None

    def zero_grad(self):
        self.ctx._check(self.lib.lux_surrogate_zero_grad(self.h), "lux_surrogate_zero_grad")

    def download(self, grads: bool = False, table: bool = True) -> dict:
        mlp = np.zeros(self.mlp_params, np.float32)
        mg = np.zeros(self.mlp_params, np.float32) if grads else None
        tb = np.zeros(self.table_params, np.float32) if table else None
        tg = np.zeros(self.table_params, np.float32) if (grads and table) else None
        self.ctx._check(self.lib.lux_surrogate_download(self.h, _ptr(mlp), _ptr(mg), _ptr(tb), _ptr(tg)),
                        "lux_surrogate_download")
        out = {"mlp": mlp}
        if grads:
            out["mlp_grad"] = mg
        if table:
            out["table"] = tb.reshape(16, -1, 2)
            if grads:
                out["table_grad"] = tg.reshape(16, -1, 2)
        return out

    def upload(self, mlp=None, table=None):
        m = None if mlp is None else np.ascontiguousarray(mlp, np.float32)
        t = None if table is None else np.ascontiguousarray(table, np.float32)
        self.ctx._check(self.lib.lux_surrogate_upload(self.h, _ptr(m), _ptr(t)), "lux_surrogate_upload")


class Batch:
    """One GI update of many configured scenes at once (lux_batch).

    Holds references to its contexts (and surrogates) so they outlive it."""

    def __init__(self, contexts, surrogates=None):
        self.contexts = list(contexts)
        if not self.contexts:
            raise ValueError("Batch needs at least one context")
        self.lib = self.contexts[0].lib
        self.surrogates = {}
        arr = (C.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        h = C.c_void_p()
        rc = self.lib.lux_batch_create(arr, len(self.contexts), C.byref(h))
        if rc != LUX_OK:
            msg = self.lib.lux_last_error(self.contexts[0].h)
            raise LuxError(f"lux_batch_create: {rc} ({msg.decode() if msg else ''})")
        self.h = h
        for i, (s, level) in (surrogates or {}).items():
            self.set_surrogate(i, s, level)

    def close(self):
        if getattr(self, "h", None):
            self.lib.lux_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != LUX_OK:
            msg = self.lib.lux_batch_last_error(self.h)
            raise LuxError(f"{what}: {rc} ({msg.decode() if msg else ''})")

    def set_surrogate(self, scene: int, surrogate, level: int = 0):
        self._check(self.lib.lux_batch_set_surrogate(self.h, scene, surrogate.h if surrogate else None, level),
                    "lux_batch_set_surrogate")
        if surrogate is None:
            self.surrogates.pop(scene, None)
        else:
            self.surrogates[scene] = (surrogate, level)

    def set_surrogate_bound(self, scene: int, error_bound: float):
        """lux_batch_set_surrogate_bound: the scene's traced per-probe fallback."""
        self._check(self.lib.lux_batch_set_surrogate_bound(self.h, scene, error_bound),
                    "lux_batch_set_surrogate_bound")
    def refresh(self):
        self._check(self.lib.lux_batch_refresh(self.h), "lux_batch_refresh")

    def set_frame(self, frame: int, seed: int):
        self._check(self.lib.lux_batch_set_frame(self.h, frame, seed), "lux_batch_set_frame")

    def update(self, timed: bool = False):
        ms = C.c_float()
        self._check(self.lib.lux_batch_update(self.h, C.byref(ms) if timed else None), "lux_batch_update")
        return float(ms.value) if timed else None

    def capture(self):
        self._check(self.lib.lux_batch_capture(self.h), "lux_batch_capture")

    def launch(self, timed: bool = False):
        ms = C.c_float()
        self._check(self.lib.lux_batch_launch(self.h, C.byref(ms) if timed else None), "lux_batch_launch")
        return float(ms.value) if timed else None

    def synchronize(self):
        self._check(self.lib.lux_batch_synchronize(self.h), "lux_batch_synchronize")

    def sizes(self) -> dict:
        it, iv, pr, sc = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint32()
        self._check(self.lib.lux_batch_sizes(self.h, C.byref(it), C.byref(iv), C.byref(pr), C.byref(sc)),
                    "lux_batch_sizes")
        return dict(items=int(it.value), intervals=int(iv.value), probes=int(pr.value), scenes=int(sc.value))

    def set_config(self, waves_per_cu: int = 24, lds_stack_depth: int = 12, regroup_threshold: int = 48,
                   steps_per_refill: int = 16):
        cfg = (C.c_uint32 * 8)(waves_per_cu, lds_stack_depth, regroup_threshold, steps_per_refill, 0, 0, 0, 0)
        self._check(self.lib.lux_batch_set_config(self.h, cfg), "lux_batch_set_config")

    def config(self) -> dict:
        cfg = (C.c_uint32 * 8)()
        self._check(self.lib.lux_batch_get_config(self.h, cfg), "lux_batch_get_config")
        return dict(waves_per_cu=cfg[0], lds_stack_depth=cfg[1], regroup_threshold=cfg[2],
                    steps_per_refill=cfg[3])

    def autotune(self, repeats: int = 2) -> dict:
        cfg = (C.c_uint32 * 8)()
        ms = C.c_float()
        ev = C.c_uint32()
        self._check(self.lib.lux_batch_autotune(self.h, repeats, cfg, C.byref(ms), C.byref(ev)),
                    "lux_batch_autotune")
        return dict(waves_per_cu=cfg[0], lds_stack_depth=cfg[1], regroup_threshold=cfg[2],
                    steps_per_refill=cfg[3], ms=float(ms.value), evaluations=int(ev.value))
