"""The Bevy-facing RadianceCascades surface through the C ABI (no GPU).

Restates the reference's src/lib.rs unit tests invalid_settings_are_bounded,
shared_sampling_and_split_low_end_preset_are_bounded and
cache_entry_limit_accepts_262144_and_clamps_larger_values (src/lib.rs:948-1009)
against lux_radiance_cascades_{default,low_end,sanitized}, and checks the
camera -> transport parameter mapping.
"""
import math

import numpy as np
import pytest

from bevy_penumbra_amd import lux


def f32(x):
    return float(np.float32(x))


def test_invalid_settings_are_bounded():
    r = lux.RadianceCascades.default()
    r.probe_spacing = math.nan
    r.resolution_divisor = 0
    r.max_rays = 0xFFFFFFFF
    r.temporal_blend = math.inf
    r.history_frames = 0xFFFFFFFF
    r.spatial_filter_steps = 0xFFFFFFFF
    r.emissive_samples = 0xFFFFFFFF
    s = r.sanitized()
    assert s.probe_spacing == 0.25
    assert s.resolution_divisor == 1
    assert s.max_rays == 131_072
    assert s.temporal_blend == f32(0.88)
    assert s.history_frames == 64
    assert s.spatial_filter_steps == 4
    assert s.emissive_samples == 32


def test_sampling_and_low_end_preset():
    d = lux.RadianceCascades.default()
    assert list(d.sampling) == [0, 0, 0, 0, 0, 1]
    assert d.resolution_divisor == 2 and d.max_rays == 131_072 and d.emissive_samples == 16
    assert d.probe_spacing == 0.25 and d.interval_length == f32(0.4)
    assert d.normal_bias == f32(0.005) and d.temporal_blend == f32(0.88)
    assert d.split_transmittance_blend == f32(0.98) and d.history_frames == 24
    assert d.flags == 16  # artistic_occlusion only
    lo = lux.RadianceCascades.low_end()
    assert list(lo.sampling) == [8, 16_384, 1_024, 65_536, 16_384, 2]
    assert lo.max_rays == 1_024 and lo.emissive_samples == 4
    assert lo.probe_spacing == 1.0 and lo.interval_length == f32(1.6)
    assert lo.history_frames == 24 and lo.temporal_blend == f32(0.95)
    assert lo.spatial_filter_steps == 3


@pytest.mark.parametrize("value,want", [(262_144, 262_144), (0xFFFFFFFF, 262_144), (0, 0)])
def test_cache_entry_limit(value, want):
    r = lux.RadianceCascades.default()
    r.sampling[3] = value
    assert r.sanitized().sampling[3] == want


def test_normal_bias_clamped_to_probe_spacing():
    r = lux.RadianceCascades.default()
    r.probe_spacing = 0.1
    r.normal_bias = 5.0
    s = r.sanitized()
    assert s.normal_bias == f32(0.1)


def test_camera_maps_onto_transport_parameters():
    r = lux.RadianceCascades.default()
    r.probe_spacing = 0.125
    r.interval_length = 0.2
    r.emissive_samples = 64  # clamped to 32
    r.history_epoch = 7
    p = r.gi_params((0, 0, 0), (16, 16, 16), 4, base_phi=4, base_y=4, interval_mode=1, frame=3)
    assert p.spacing == 0.125 and p.interval_length == f32(0.2)
    assert p.emissive_samples == 32 and p.seed == 7 and p.frame == 3
    assert list(p.dims) == [16, 16, 16] and p.levels == 4 and p.terminal_tail == 1


def test_traversal_query_budget_reaches_transport_parameters():
    r = lux.RadianceCascades.default()
    p = r.gi_params((0, 0, 0), (8, 8, 8), 3)
    assert p.query_limit == 0                  # unbounded by default
    low = lux.RadianceCascades.low_end()
    assert low.sampling[4] == 16384
    assert low.gi_params((0, 0, 0), (8, 8, 8), 3).query_limit == 16384
    low.sampling[4] = 10 ** 7                  # sanitized into 1..131072
    assert low.gi_params((0, 0, 0), (8, 8, 8), 3).query_limit == 131072
