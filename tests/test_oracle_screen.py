"""Oracle of the screen-space temporal and denoise passes.

Pinned on the reference's own fixture (src/transport_tests.rs:1235-1530 with
its textures at :440-630 and the denoise case at :854-1089), restated scenario
by scenario: unobserved history retained without gaining age, a black
observation after a lighting revision, quarter-coverage raw alpha with
history off, dirty bounds disjoint from / containing the receiver, lighting
revisions with incomplete and partial refreshes (and the frame after), a
gradually moving sun converging once it stops, an abrupt extinction, expired
unsampled sun history, moving punctual lights keeping the separate
background, the clamped last resolve tile at scales 2..8, and the denoise
filling an unsupported centre from its own receiver but not another's (flat
and curved).  Beyond those, invariants: with history reuse off a supported
sample passes through unchanged; denoising a constant field returns it.
"""
import pytest

from tests import screen_fixture as F

B = F.OracleBackend


def test_unobserved_history_black_observation_and_partial_coverage():
    F.check_unobserved_history_black_observation_and_partial_coverage(B)


def test_dirty_bounds_keep_or_reject_history():
    F.check_dirty_bounds_keep_or_reject_history(B)


@pytest.mark.parametrize("alpha,length,mask", [(-1.0, 0.0, 0), (1.0625, 0.25, 1)])
def test_lighting_revision_discards_old_radiance(alpha, length, mask):
    F.check_lighting_revision_discards_old_radiance(B, alpha, length, mask)


def test_gradual_sun_blends_converges_and_expires():
    F.check_gradual_sun_blends_converges_and_expires(B)


@pytest.mark.parametrize("alpha,current,expected", [(1.25, 12.0, 21.0), (-1.0, 10.0, 19.0)])
def test_moving_punctual_lights_keep_the_background(alpha, current, expected):
    F.check_moving_punctual_lights_keep_the_background(B, alpha, current, expected)


@pytest.mark.parametrize("scale", range(2, 9))
def test_static_reprojection_reads_the_clamped_last_tile(scale):
    F.check_static_reprojection_reads_the_clamped_last_tile(B, scale)


@pytest.mark.parametrize("curved", [False, True])
def test_denoise_fills_from_its_own_receiver_only(curved):
    F.check_denoise_fills_from_its_own_receiver_only(B, curved)


def test_history_off_passes_a_supported_sample_through():
    F.check_history_off_passes_a_supported_sample_through(B)


def test_denoise_of_a_constant_field_returns_it():
    F.check_denoise_of_a_constant_field_returns_it(B)


def test_packed_normal_matches_the_fixture_words():
    F.check_packed_normal_matches_the_fixture_words(B)
