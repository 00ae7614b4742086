"""Oracle traversal against the f64 brute force on the reference's fixtures.

Scene: 24 instances (cube/panel/bare, mirrored and extreme scales, cull modes)
as in src/traversal_tests.rs:2120-2177; rays: src/traversal_tests.rs:1040-1125.
Forced deep chains must report error bits 1 (BLAS) and 2 (TLAS) on hits and
still return every hit (src/traversal_tests.rs:2270-2283).
"""
import numpy as np
import pytest

from bevy_penumbra_amd import scenes
from oracle import oracle as O
from tests.traversal_check import check_against_brute_force, forced_deep, first_hit_copies


@pytest.fixture(scope="module")
def fixture_scene():
    sc = scenes.traversal_fixture()
    rays = scenes.fixture_rays(sc)
    return sc, rays


@pytest.mark.parametrize("sah", [False, True])
def test_fixture_traversal_matches_brute_force(fixture_scene, sah):
    sc, rays = fixture_scene
    built = O.build_scene(sc, sah=sah)
    all_rays = np.concatenate([rays, first_hit_copies(rays[rays.shape[0] // 2:])])
    hits = built.trace(all_rays)
    failures, counts = check_against_brute_force(built, all_rays, hits)
    assert not failures, "\n".join(failures)
    assert counts["closest"] > 100 and counts["shadow"] > 100 and counts["first"] > 100, counts


@pytest.mark.parametrize("blas", [True, False])
def test_forced_deep_chains_report_overflow(fixture_scene, blas):
    sc, rays = fixture_scene
    built = O.build_scene(sc)
    deep = forced_deep(built, blas)
    sel = np.concatenate([rays[:1500], rays[rays.shape[0] // 2: rays.shape[0] // 2 + 1500]])
    sel = np.concatenate([sel, first_hit_copies(sel[1500:])])
    hits = deep.trace(sel)
    failures, counts = check_against_brute_force(deep, sel, hits, expected_errors=1 if blas else 2)
    assert not failures, "\n".join(failures)
    assert counts["closest"] > 20 and counts["shadow"] > 20


def test_lbvh_and_sah_give_the_same_closest_hits(fixture_scene):
    sc, rays = fixture_scene
    a = O.build_scene(sc, sah=False).trace(rays)
    b = O.build_scene(sc, sah=True).trace(rays)
    closest = rays.view(np.uint32)[:, 7] == 0
    same = np.all(a[closest][:, :3] == b[closest][:, :3], axis=1)
    # topology changes the visit order; only exact-distance ties may differ
    assert same.mean() > 0.95
    ta = a[closest][~same, 0].view(np.float32)
    tb = b[closest][~same, 0].view(np.float32)
    assert np.all(ta == tb)
    blocked_a = (a[~closest, 3] & 0x100) != 0
    blocked_b = (b[~closest, 3] & 0x100) != 0
    assert np.array_equal(blocked_a, blocked_b)


def test_colonnade_sample_matches_brute_force():
    sc, _ = scenes.colonnade(triangles=262_144 // 16)
    built = O.build_scene(sc)
    rays = scenes.microbench_rays(sc.bounds, n=400, seed=5)
    hits = built.trace(rays)
    failures, counts = check_against_brute_force(built, rays, hits)
    assert not failures, "\n".join(failures)
    assert counts["closest"] > 50 and counts["shadow"] > 50


def test_empty_ray_set_and_miss_outside():
    sc = scenes.traversal_fixture()
    built = O.build_scene(sc)
    assert built.trace(np.zeros((0, 8), np.float32)).shape == (0, 4)
    r = np.array([[500, 500, 500, 10, 1, 0, 0, 0]], np.float32)
    h = built.trace(r)
    assert h[0, 1] == 0xFFFFFFFF and h[0, 3] == 0
