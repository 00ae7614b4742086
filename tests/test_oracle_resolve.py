"""Oracle of the split resolve (LOD pair) against the reference's fixtures.

The C oracle's resolve over dense LOD grids is checked against the
reference's own CPU model of the resolve on the reference's fixture cases
(tests/resolve_fixtures.py, src/world_resolve_tests.rs:749-980) within the
reference's tolerance, and its LOD choice against values worked out by hand
from splitLodBlendForParameters (shaders/common.slang:431-461).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import resolve_fixtures as F


def grid_params(lod: int) -> O.GiParams:
    return O.gi_params(F.GRID_ORIGIN, float(2 ** lod), F.grid_dims(lod), 1)


def resolve_fixture(probes, perspective, receivers, lods=3):
    view = O.lod_view(**(F.perspective_view_terms() if perspective else F.orthographic_view_terms()))
    grids = [grid_params(k) for k in range(lods)]
    return O.world_resolve_lods(grids, F.dense_tables(probes, lods), view, receivers)


@pytest.mark.parametrize("case", [c[0] for c in F.fixture_cases()])
def test_oracle_resolve_matches_reference_fixture(case):
    name, probes, perspective, rec = next(c for c in F.fixture_cases() if c[0] == case)
    out = resolve_fixture(probes, perspective, rec)
    ref_probes = F.f32_tables(probes)
    for r, o in zip(rec, out):
        if name == "angular":
            table = ref_probes[next(iter(ref_probes))]
            val, cov = F.angular_lookup(table, r[3:])
        else:
            val, cov = F.split_diffuse(ref_probes, r[:3].astype(np.float64), r[3:].astype(np.float64),
                                       perspective)
        for a, e in zip(o[:3], val):
            assert F.close(float(a), float(e)), (name, o, val, cov)
        assert F.close(float(o[3]), F.expected_alpha(cov)), (name, o, cov)
    if name == "sparse_support":
        assert np.all(out[3, :3] == 0.0) and out[3, 3] > 1.0     # supported black stays supported
        assert out[4, 3] == -1.0                                 # missing probes stay unsupported
    if name == "partial_black":
        assert np.all(out[0, :3] == 0.0) and 1.0 < out[0, 3] < 1.25
    if name == "tiny_support":
        assert np.all(out[0, :3] > 0.0) and out[0, 3] > 0.0
    if name == "angular":
        assert np.all(np.abs(out[2, :3] - out[3, :3]) < 0.02)   # octahedral south-pole seam
    if name == "zero_coverage":
        assert out[0].tolist() == [0.0, 0.0, 0.0, -1.0]


def test_oracle_overlap_blends_half_and_half():
    """At z = -1.9 the perspective fixture camera gives LOD scale 1.9: low
    LOD 0, high LOD 1, blend (1.9 - 1.8) / 0.2 = 0.5."""
    lo, hi, b = O.lod_select(O.lod_view(**F.perspective_view_terms()), (0.0, 0.0, -1.9))
    assert (lo, hi) == (0, 1) and abs(b - 0.5) < 1e-5


# LOD scale s -> (low, high, blend), by hand: low = floor(log2 max(s, 1)),
# overlap [0.9 * 2^(low+1), 2^(low+1)) blends linearly toward low + 1
HAND = [(0.5, 0, 0, 0.0), (1.0, 0, 0, 0.0), (1.5, 0, 0, 0.0), (1.8, 0, 1, 0.0), (1.9, 0, 1, 0.5),
        (2.0, 1, 1, 0.0), (3.7, 1, 2, 0.25), (7.5, 2, 3, 0.375), (8.0, 3, 3, 0.0), (30.0, 4, 5, 0.375),
        (31.0, 4, 5, 0.6875), (2.0 ** 25, 24, 24, 0.0), (2.0 ** 30, 24, 24, 0.0)]


@pytest.mark.parametrize("s,low,high,blend", HAND)
def test_oracle_lod_choice_by_hand(s, low, high, blend):
    view = O.lod_view(**F.perspective_view_terms())
    lo, hi, b = O.lod_select(view, (s, 0.25, -0.25))
    assert (lo, hi) == (low, high) and abs(b - blend) < 1e-4, (s, lo, hi, b)
    # translation: the same offset from a moved camera
    moved = O.lod_view(**dict(F.perspective_view_terms(), camera=(10.0, -3.0, 2.0)))
    lo2, hi2, b2 = O.lod_select(moved, (10.0 - s, -2.75, 1.75))
    assert (lo2, hi2) == (lo, hi) and abs(b2 - b) < 1e-4


def test_oracle_lod_base_distance_and_orthographic():
    """Perspective: base distance max(focal_px * spacing / target, spacing)
    (focal_px = 0.5 * height * |row1|); orthographic: scale = target /
    (focal_px * spacing)."""
    # focal 0.5 * 200 * 1 = 100 px, spacing 0.5, 2 target px: base 25 m
    v = O.lod_view(camera=(0, 0, 0), probe_spacing=0.5, clip_row1=(0, 1, 0), viewport_height=200,
                   clip_row3=(0, 0, -1), target_pixels=2)
    assert O.lod_select(v, (92.5, 0, 0)) == (1, 2, pytest.approx(0.25, abs=1e-5))  # s = 3.7
    assert O.lod_select(v, (49.0, 0, 0))[:2] == (0, 1)                             # s = 1.96
    # orthographic: focal 0.5 * 2 * 0.25 = 0.25 px, spacing 0.5, 15 px: s = 120
    o = O.lod_view(camera=(0, 0, 0), probe_spacing=0.5, clip_row1=(0, 0.25, 0), viewport_height=2,
                   clip_row3=(0, 0, 0), target_pixels=15)
    lo, hi, b = O.lod_select(o, (1e6, 0, 0))
    assert (lo, hi) == (6, 7) and abs(b - (120 - 115.2) / 12.8) < 1e-4
    # target 0 counts as 1; a non-finite row 3 is not a perspective camera
    z = O.lod_view(camera=(0, 0, 0), probe_spacing=1.0, clip_row1=(0, 1, 0), viewport_height=2,
                   clip_row3=(float("nan"), 0, -1), target_pixels=0)
    assert O.lod_select(z, (5.0, 0, 0)) == (0, 0, 0.0)


def test_oracle_lod_pyramid_reveals_the_blend():
    """Uniform LOD grids of value 10^k (support 1): the resolved value is
    10^low (1 - blend) + 10^high blend, pinning choice and mix together."""
    view = O.lod_view(**F.perspective_view_terms())
    grids, tables = [], []
    for k in range(4):
        g = O.gi_params((-16.0, -16.0, -16.0), float(2 ** k), (32 >> k,) * 3, 1)
        t = np.zeros((g.dims[0] ** 3, 36, 4), np.float32)
        t[..., :3] = 10.0 ** k
        t[..., 3] = 1.0
        grids.append(g)
        tables.append(t)
    for s, low, high, blend in HAND[:8]:
        rec = np.array([[s, 0.25, 0.25, 0.0, 1.0, 0.0]], np.float32)
        out = O.world_resolve_lods(grids, tables, view, rec)[0]
        want = 10.0 ** low * (1 - blend) + 10.0 ** high * blend
        assert abs(out[0] - want) <= 1e-5 * want + 2e-4 * 10.0 ** high * (blend != 0), (s, out, want)
        assert out[3] == 1.25


def test_single_lod_resolve_is_the_lod_pyramid_of_one():
    """lux_world_resolve_device is the pyramid of one grid under a camera
    that always picks LOD 0 (orthographic, scale 1)."""
    p = grid_params(0)
    pr = {}
    F.insert_probes(pr, (0, 0, 0), 0, [2.0, 1.0, 3.0], 0xA55AC33C, 0x5F)
    tables = F.dense_tables(pr, 1)
    rng = np.random.default_rng(3)
    rec = np.concatenate([rng.uniform(-1.5, 1.5, (400, 3)), rng.normal(size=(400, 3))], axis=1).astype(np.float32)
    a = O.world_resolve(p, tables[0], rec)
    b = O.world_resolve_lods([p], tables, O.lod_view(**F.orthographic_view_terms()), rec)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
