"""The dense interval trace pinned on the reference's dense 16^3 fixture.

src/transport_tests.rs:3830-4106 traces a 16^3 grid of probes at spacing 1
from -8 over 4 levels with 32 * 4^L equal-area directions and intervals
[2^L - 1, 2^(L+1) - 1] through a scene of a box, a sphere, an emissive slab
(6, 2, 0.5) and a thin blocker, expecting J = emission when the first surface
inside the interval is the slab, J = 0 on other surfaces, and beta = 1 when
the interval is empty.  The same grid runs here through the oracle's
interval trace (the sphere as a box, every shape as triangles) and every
interval is checked against f64 slab intersections of its segment -- for
both interval laws and with the terminal tail on and off -- so the level
segmentation (segment origin c + d * start, tmax = end - start, the terminal
tail) and the interval laws are fixed by something outside the oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.transport_cases import (DENSE_BOXES, DENSE_EMISSION, dense_expected, dense_fixture_params,
                                   dense_fixture_scene, dense_interval_laws)


def test_interval_bounds_closed_forms():
    small = O.gi_params((0, 0, 0), 0.125, (16, 16, 16), 4, interval_mode=1, interval_length=0.2)
    want = [(0.0, 0.2), (0.2, 0.6), (0.6, 1.4), (1.4, 3.0)]
    for L, (s, e) in enumerate(want):
        got = O.interval_bounds(small, L)
        assert abs(got[0] - s) < 1e-6 and abs(got[1] - e) < 1e-6, (L, got)
    fixture = dense_fixture_params(interval_mode=1)
    for L in range(4):
        assert O.interval_bounds(fixture, L) == (float(2 ** L - 1), float(2 ** (L + 1) - 1))
    prod = O.gi_params((0, 0, 0), 0.5, (8, 8, 8), 6, interval_mode=0, interval_length=0.5)
    for L in range(6):
        assert O.interval_bounds(prod, L) == (0.0 if L == 0 else 0.5 * 4.0 ** (L - 1), 0.5 * 4.0 ** L)
    for mode, length in ((1, 0.2), (0, 0.5), (1, 1.0)):
        p = O.gi_params((0, 0, 0), 1.0, (4, 4, 4), 5, interval_mode=mode, interval_length=length)
        for L, (s, e) in enumerate(dense_interval_laws(mode, length, 5)):
            got = O.interval_bounds(p, L)
            assert np.isclose(got[0], s, rtol=1e-6) and np.isclose(got[1], e, rtol=1e-6)


def test_dense_fixture_layout():
    p = dense_fixture_params()
    assert O.total_intervals(p) == 245_760                   # RECORDS of the reference fixture
    for L in range(4):
        n, dims = O.level_dims(p, L)
        assert dims == (16 >> L,) * 3 and O.level_directions(p, L) == 32 << (2 * L)


@pytest.fixture(scope="module")
def dense_built():
    sc = dense_fixture_scene()
    return sc, O.build_scene(sc), O.build_lights(sc)


@pytest.mark.parametrize("mode,terminal", [(1, 0), (1, 1), (0, 0), (0, 1)])
def test_dense_fixture_intervals_match_slab_intersections(dense_built, mode, terminal):
    sc, built, lights = dense_built
    p = dense_fixture_params(interval_mode=mode, terminal_tail=terminal)
    iv = O.gi_trace(built, lights, sc.materials, p, threads=8)
    want, close, mats = dense_expected(p)
    assert iv.shape == want.shape
    assert close.mean() < 0.01, close.mean()
    for b in range(len(DENSE_BOXES)):                        # every shape is hit, as the reference asserts
        assert np.any(mats[~close] == b), b
    ok = ~close
    assert np.array_equal(iv[ok, 3], want[ok, 3].astype(np.float32)), np.nonzero(iv[ok, 3] != want[ok, 3])[0][:5]
    assert np.array_equal(iv[ok, :3], want[ok, :3].astype(np.float32))
    emissive = ok & (mats == 2)                              # the slab
    assert emissive.sum() > 100 and np.all(iv[emissive, :3] == np.float32(DENSE_EMISSION))
    # levels whose interval starts inside the scene hold hits (mode 0's level 3
    # starts 16 m out, past every shape), and without the tail empty intervals
    for L in range(4):
        a, b = O.level_offset(p, L), O.level_offset(p, L + 1)
        if mode == 1 or L < 3:
            assert np.any(iv[a:b, 3] == 0.0), L
        if not (terminal and L == 3):
            assert np.any(iv[a:b, 3] == 1.0), L


def test_dense_fixture_fp64_twin_agrees(dense_built):
    sc, built, lights = dense_built
    p = dense_fixture_params()
    iv = O.gi_trace(built, lights, sc.materials, p, threads=8)
    iv64 = O.gi_trace(built, lights, sc.materials, p, threads=8, f64=True)
    _, close, _ = dense_expected(p)
    assert np.array_equal(iv[~close], iv64[~close].astype(np.float32))
