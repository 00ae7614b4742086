"""The temporal / denoise fixture of src/transport_tests.rs:1235-1530
(textures :440-630, denoise case :854-1089), restated as scenario checks
that run on a backend: the oracle (tests/test_oracle_screen.py) or the HIP
passes (tests/test_gpu_screen.py).  Textures.backend selects it."""

import ctypes as C
import struct

import numpy as np

from oracle import oracle as O

IDENTITY = np.eye(4, dtype=np.float32).ravel()


def fbits(f):
    return struct.unpack("<I", struct.pack("<f", f))[0]


def params(history_frames, width=1):
    """temporal_test_parameters (src/transport_tests.rs:1090-1125)."""
    P = O.ScreenParams()
    for m in ("world_from_clip", "previous_clip_from_world", "previous_world_from_clip"):
        getattr(P, m)[:] = IDENTITY
    P.viewport[:] = [0, 0, width, 1]
    P.target[:] = [width, 1]
    P.size[:] = [1, 1]
    P.resolve_scale = 0
    P.persistent = 0
    P.probe_spacing, P.normal_bias = 1.2, 0.005
    P.history_blend, P.history_frames = 0.9, float(history_frames)
    P.stability[:] = [1.0, 0.0, 0.0, 0.0]
    return P


class OracleBackend:
    temporal = staticmethod(O.screen_temporal)
    denoise = staticmethod(O.screen_denoise)


class Textures:
    """TemporalTestTextures::new_sized: depth 0.5, previous (1, 1, 1, 0.5),
    previous metadata (0.5, 0.5, 16, 0), previous receiver (1, 2, 8390656,
    65537), raster identity (1.0, 2.0) bits, G-buffer normal word 0x00800800;
    the rest zero.  `backend.temporal` runs the pass."""

    def __init__(self, n=1, backend=OracleBackend):
        self.n = n
        self.backend = backend
        z = lambda: np.zeros((1, n, 4), np.float32)
        self.depth = np.full((1, n), 0.5, np.float32)
        self.current, self.current_background, self.previous_background = z(), z(), z()
        self.previous = np.tile(np.float32([1, 1, 1, 0.5]), (1, n, 1))
        self.previous_metadata = np.tile(np.float32([0.5, 0.5, 16, 0]), (1, n, 1))
        self.previous_receiver = np.tile(np.float32([1, 2, 8_390_656, 65_537]), (1, n, 1))
        self.motion = np.zeros((1, n, 2), np.float32)
        self.raster = np.tile(np.uint32([fbits(1.0), fbits(2.0), 0, 0]), (1, n, 1))
        self.gbuffer = np.tile(np.uint32([0, 0, 0, 0x0080_0800]), (1, n, 1))

    def run(self, P, advance=False):
        RW, RH = O.screen_extent(P)
        r = lambda a: a[:RH, :RW]
        out = self.backend.temporal(P, r(self.current), r(self.current_background), self.depth, self.gbuffer,
                                self.raster, self.motion, r(self.previous), r(self.previous_metadata),
                                r(self.previous_receiver), r(self.previous_background))
        if advance:
            for name, o in zip(("previous", "previous_metadata", "previous_receiver", "previous_background"), out):
                getattr(self, name)[:RH, :RW] = o
        return out


def support(receiver):
    return (int(receiver[3]) >> 16) & 1


def px(out, i=0):
    """(irradiance, metadata, receiver) of resolve texel i."""
    return out[0][0, i], out[1][0, i], out[2][0, i]


def check_unobserved_history_black_observation_and_partial_coverage(backend):
    t = Textures(backend=backend)
    t.current[0, 0] = [0, 0, 0, -1]
    P = params(30)
    P.persistent = 1
    irr, meta, recv = px(t.run(P))
    assert np.allclose(irr[:3], 1.0, atol=1e-4) and abs(irr[3] - 0.5) < 1e-5
    assert abs(meta[2] - 16.0) < 1e-5 and support(recv) == 1
    t = Textures(backend=backend)
    t.current[0, 0] = [0, 0, 0, 1.25]
    P = params(30)
    P.persistent, P.stability[1], P.source_changes = 1, 1.0, 7
    irr, meta, recv = px(t.run(P))
    assert np.abs(irr[:3]).max() < 1e-6 and abs(meta[2] - 1.0) < 1e-5 and support(recv) == 1
    t = Textures(backend=backend)
    t.current[0, 0] = [0.25, 0.25, 0.25, 1.0625]
    P = params(30)
    P.persistent, P.history_blend = 1, 0.0
    irr, meta, recv = px(t.run(P))
    assert np.allclose(irr[:3], 0.25, atol=1e-5) and abs(meta[2] - 0.25) < 1e-5 and support(recv) == 1


def check_dirty_bounds_keep_or_reject_history(backend):
    t = Textures(backend=backend)
    t.current[0, 0] = [0, 0, 0, -1]
    P = params(30)
    P.persistent, P.stability[1] = 1, 1.0
    P.dirty_min[:], P.dirty_max[:] = [10, 10, 10, 1], [11, 11, 11, 1]
    irr, meta, recv = px(t.run(P))
    assert np.allclose(irr[:3], 1.0, atol=1e-4) and abs(meta[2] - 16.0) < 1e-5 and support(recv) == 1
    P.dirty_min[:], P.dirty_max[:] = [0, -2, 0, 1], [2, 0, 1, 1]
    irr, meta, recv = px(t.run(P))
    assert np.abs(irr[:3]).max() < 1e-6 and irr[3] == 0.0 and support(recv) == 0


def check_lighting_revision_discards_old_radiance(backend, alpha, length, mask):
    t = Textures(backend=backend)
    t.current[0, 0] = [0, 0, 0, alpha]
    P = params(30)
    P.persistent, P.stability[1], P.source_changes = 1, 1.0, 7
    irr, meta, recv = px(t.run(P, advance=True))
    assert np.abs(irr[:3]).max() < 1e-6 and abs(meta[2] - length) < 1e-5 and support(recv) == mask
    P.stability[1], P.source_changes = 0.0, 0
    t.current[0, 0] = [0, 0, 0, -1]
    irr, meta, recv = px(t.run(P))
    assert np.abs(irr[:3]).max() < 1e-6 and support(recv) == mask


def check_gradual_sun_blends_converges_and_expires(backend):
    t = Textures(backend=backend)
    t.current[0, 0] = [0, 0, 0, 1.25]
    t.current_background[0, 0] = [0, 0, 0, 1.25]
    P = params(30)
    P.persistent, P.source_changes, P.source_flags = 1, 1, 3
    last = 1.0
    for frame in range(30, 50):
        P.frame, P.motion_epoch = frame, frame
        irr, _, recv = px(t.run(P, advance=True))
        assert support(recv) == 1 and 0.0 < irr[0] < last
        last = irr[0]
    P.source_changes, P.source_flags = 0, 2
    for frame in range(50, 90):
        P.frame = frame
        last = px(t.run(P, advance=True))[0][0]
    assert last < 0.05
    e = Textures(backend=backend)
    e.current[0, 0] = [0, 0, 0, 1.25]
    P.source_changes, P.source_flags = 1, 2
    assert np.abs(px(e.run(P))[0][:3]).max() < 1e-6                # abrupt extinction
    u = Textures(backend=backend)
    u.current[0, 0] = [0, 0, 0, -1]
    P.source_changes, P.source_flags, P.motion_epoch, P.frame = 0, 2, 49, 170
    irr, _, recv = px(u.run(P))
    assert support(recv) == 0 and irr[3] == 0.0                      # unsampled old sun expired


def check_moving_punctual_lights_keep_the_background(backend, alpha, current, expected):
    t = Textures(backend=backend)
    t.current[0, 0] = [current] * 3 + [alpha]
    t.current_background[0, 0] = [10, 10, 10, 1.25]
    t.previous_background[0, 0] = [20, 20, 20, 16]
    P = params(30)
    P.persistent, P.source_changes = 1, 4
    irr, _, recv = px(t.run(P))
    assert np.allclose(irr[:3], expected, atol=1e-4) and support(recv) == 1
    P.source_changes = 5
    assert np.allclose(px(t.run(P))[0][:3], current, atol=1e-4)
    P.source_changes = 4
    P.dirty_min[:], P.dirty_max[:] = [10, 10, 10, 1], [11, 11, 11, 1]
    assert np.allclose(px(t.run(P))[0][:3], current, atol=1e-4)


def check_static_reprojection_reads_the_clamped_last_tile(backend, scale):
    n = scale + 1
    t = Textures(n, backend)
    t.previous[0, 1] = [5, 5, 5, 0.5]
    t.current[0, :] = [0, 0, 0, -1]
    P = params(30, n)
    P.persistent, P.resolve_scale = 1, scale
    assert O.screen_extent(P) == (2, 1)
    irr, _, recv = px(t.run(P), 1)
    assert np.allclose(irr[:3], 5.0, atol=1e-5) and support(recv) == 1


def denoise_case(curved):
    """assert_denoise_rejects_other_receiver (src/transport_tests.rs:854-1089)."""
    n = 11

    def receiver(identity, supported):
        return [struct.unpack("<f", struct.pack("<I", identity[0]))[0],
                struct.unpack("<f", struct.pack("<I", identity[1]))[0],
                8_390_656.0, 32_769.0 + (65_536.0 if supported else 0.0)]

    inp = np.zeros((1, n, 4), np.float32)
    mom = np.zeros((1, n, 4), np.float32)
    rec = np.tile(np.float32(receiver([90, 91], False)), (1, n, 1))
    inp[0, 0], inp[0, 1], inp[0, 2] = [1, 0, 0, 0.5], [0, 1, 0, 0], [0, 0, 1, 0.5]
    mom[0, 0] = mom[0, 2] = [0, 0, 1, 1]
    rec[0, 0] = receiver([1, 2], True)
    rec[0, 1] = receiver([1, 2], False)
    rec[0, 2] = receiver([3, 4], True)
    if curved:
        def packed(angle):
            x = np.sin(np.float32(angle)) / (abs(np.sin(np.float32(angle))) + np.cos(np.float32(angle)))
            return float(int(round((x * 0.5 + 0.5) * 4095.0)) | (2048 << 12))
        rec[0, 0, 2] = packed(-0.1)
        rec[0, 1, 2] = packed(0.1)
    P = params(64, n)
    P.probe_spacing = 0.15
    P.size[:] = [n, 1]
    depth = np.full((1, n), 0.5, np.float32)
    return P, inp, mom, rec, depth


def check_denoise_fills_from_its_own_receiver_only(backend, curved):
    P, inp, mom, rec, depth = denoise_case(curved)
    out, out_r = backend.denoise(P, 0, inp, mom, rec, depth)
    hole = out[0, 1]
    assert hole[0] > 0.9 and hole[1] < 1e-5 and hole[2] < 1e-5 and abs(hole[3] - 0.5) < 1e-5
    assert support(out_r[0, 1]) == 1 and support(out_r[0, 10]) == 0


def check_history_off_passes_a_supported_sample_through(backend):
    rng = np.random.default_rng(4)
    t = Textures(7, backend)
    t.current[0, :, :3] = rng.uniform(0, 5, (7, 3))
    t.current[0, :, 3] = 1.0                                          # supported, full freshness
    P = params(30, 7)
    P.stability[0] = 0.0                                              # no history reuse
    irr, meta, recv, _ = t.run(P)
    assert np.array_equal(irr[0, :, :3], t.current[0, :, :3]) and (meta[0, :, 2] == 1.0).all()
    assert all(support(r) == 1 for r in recv[0])


def check_denoise_of_a_constant_field_returns_it(backend):
    n = 9
    P = params(30, n)
    P.size[:] = [n, 1]
    inp = np.tile(np.float32([0.3, 0.6, 0.9, 0.5]), (1, n, 1))
    mom = np.tile(np.float32([0, 0, 1, 0.01]), (1, n, 1))
    rec = np.tile(np.float32([1, 2, 8_390_656, 65_537]), (1, n, 1))
    for p in range(3):
        out, out_r = backend.denoise(P, p, inp, mom, rec, np.full((1, n), 0.5, np.float32))
        assert np.allclose(out[0, :, :3], [0.3, 0.6, 0.9], rtol=1e-6) and (out[0, :, 3] == 0.5).all()


def check_packed_normal_matches_the_fixture_words(backend):
    # the G-buffer word 0x00800800 decodes to +z, which packs to 2048 | 2048 << 12
    assert O.screen_packed_normal([0.0, 0.0, 1.0]) == 8_390_656.0
    assert C.sizeof(O.ScreenParams) == 312
