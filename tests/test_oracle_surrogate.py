"""Pins of the surrogate oracle (oracle/surrogate.py) by mathematics.

The reference has no learned component, so nothing here is a reference
vector: each check is a closed form, an independent library routine or a
finite difference chosen so that a plausible slip (a dropped phase, a
swapped index, a transposed weight, a missing ReLU mask) fails it.
"""
import math

import numpy as np
import pytest

from oracle import surrogate as S


def unit_dirs(n, seed=0):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


# ------------------------------------------------------------------ SH
def scipy_real_sh(d):
    """Real SH from scipy's complex Y_l^m (which carries the Condon-Shortley
    phase): sqrt(2) (-1)^m Re/Im Y_l^|m|, polar axis z."""
    sp = pytest.importorskip("scipy.special")
    theta = np.arccos(np.clip(d[:, 2], -1, 1))
    phi = np.arctan2(d[:, 1], d[:, 0])
    out = np.zeros((d.shape[0], 64))
    for l in range(8):
        for m in range(0, l + 1):
            y = sp.sph_harm_y(l, m, theta, phi)
            if m == 0:
                out[:, l * l + l] = y.real
            else:
                out[:, l * l + l + m] = math.sqrt(2) * (-1) ** m * y.real
                out[:, l * l + l - m] = math.sqrt(2) * (-1) ** m * y.imag
    return out


def test_sh_matches_scipy():
    d = unit_dirs(500)
    ref = scipy_real_sh(d)
    np.testing.assert_allclose(S.sh_encode(d, exact=False), ref, atol=1e-12, rtol=1e-10)
    # the f32 recurrence the GPU runs stays within f32 rounding of the f64 values
    np.testing.assert_allclose(S.sh_encode(d.astype(np.float32), exact=True), ref, atol=2e-5)


def test_sh_orthonormal_by_quadrature():
    # Gauss-Legendre in z (16 nodes: exact to degree 31) x 32 uniform phi
    # nodes (exact for trigonometric degree < 32) integrate every product of
    # two band <= 7 harmonics exactly.
    z, wz = np.polynomial.legendre.leggauss(16)
    phi = (np.arange(32) + 0.5) * (2 * np.pi / 32)
    Z, P = np.meshgrid(z, phi, indexing="ij")
    r = np.sqrt(1 - Z ** 2)
    d = np.stack([(r * np.cos(P)).ravel(), (r * np.sin(P)).ravel(), Z.ravel()], axis=1)
    w = (wz[:, None] * np.full(32, 2 * np.pi / 32)[None, :]).ravel()
    Y = S.sh_encode(d, exact=False)
    gram = (Y * w[:, None]).T @ Y
    np.testing.assert_allclose(gram, np.eye(64), atol=1e-12)


def test_sh_band_zero_and_one():
    d = unit_dirs(10, 3)
    Y = S.sh_encode(d, exact=False)
    np.testing.assert_allclose(Y[:, 0], 0.5 / math.sqrt(math.pi))
    c1 = math.sqrt(3 / (4 * math.pi))
    np.testing.assert_allclose(Y[:, 1], c1 * d[:, 1], atol=1e-14)  # m = -1 -> y
    np.testing.assert_allclose(Y[:, 2], c1 * d[:, 2], atol=1e-14)  # m = 0 -> z
    np.testing.assert_allclose(Y[:, 3], c1 * d[:, 0], atol=1e-14)  # m = +1 -> x


# ------------------------------------------------------------------ grid
def test_level_resolutions_and_dense_levels():
    res, dense = S.level_resolutions(16.0, 1.5)
    assert res[:6] == [16, 24, 36, 54, 81, 121]
    assert res[15] == math.floor(16 * 1.5 ** 15)
    assert dense == [True] * 4 + [False] * 12      # 55^3 <= 2^19 < 82^3


def test_grid_index_closed_forms():
    r = 16
    x, y, z = np.meshgrid(np.arange(r + 1), np.arange(r + 1), np.arange(r + 1), indexing="ij")
    idx = S.grid_index(x.ravel(), y.ravel(), z.ravel(), r, True)
    assert len(np.unique(idx)) == (r + 1) ** 3 and idx.max() == (r + 1) ** 3 - 1
    assert S.grid_index(3, 5, 7, r, True) == 3 + 17 * (5 + 17 * 7)
    h = S.grid_index(np.array([1, 0, 0, 1000]), np.array([1, 1, 0, 2000]), np.array([1, 0, 1, 3000]), 4670, False)
    exp = [((1 ^ 2654435761 ^ 805459861) & 0xFFFFFFFF) & (S.TABLE - 1),
           2654435761 & (S.TABLE - 1),
           805459861 & (S.TABLE - 1),
           ((1000 ^ ((2000 * 2654435761) & 0xFFFFFFFF) ^ ((3000 * 805459861) & 0xFFFFFFFF))) & (S.TABLE - 1)]
    assert h.tolist() == exp


def test_trilinear_reproduces_linear_fields():
    enc = S.Encoder(np.zeros(3), np.full(3, 4.0))
    rng = np.random.default_rng(1)
    pos = rng.uniform(0, 4, size=(300, 3)).astype(np.float32)
    table = np.zeros((S.LEVELS, S.TABLE, 2))
    for l in range(4):                          # dense levels: field = f(grid coords)
        r = enc.res[l]
        g = np.arange(r + 1)
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        idx = S.grid_index(X.ravel(), Y.ravel(), Z.ravel(), r, True)
        table[l, idx, 0] = 1.0 * X.ravel() + 2.0 * Y.ravel() - 0.5 * Z.ravel()
        table[l, idx, 1] = 7.0
    feats = enc.hash_features(pos, table, exact=False)
    p = enc.unit_position(pos).astype(np.float64)
    for l in range(4):
        gp = p * np.float32(enc.res[l])
        np.testing.assert_allclose(feats[:, 2 * l], gp[:, 0] + 2 * gp[:, 1] - 0.5 * gp[:, 2], rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(feats[:, 2 * l + 1], 7.0, rtol=1e-6)
    # the eight weights of every level sum to one
    for l in (0, 7, 15):
        wsum = sum(w.astype(np.float64) for _, w in enc.corners(pos, l))
        np.testing.assert_allclose(wsum, 1.0, atol=1e-6)


def test_positions_are_clamped_to_the_box():
    enc = S.Encoder(np.array([-1, -1, -1]), np.array([1, 1, 1]))
    p = enc.unit_position(np.array([[-5, 0, 5], [1, 1, 1], [np.nan, np.inf, -np.inf]], np.float32))
    assert p[0, 0] == 0 and p[0, 1] == np.float32(0.5) and p[0, 2] == np.float32(0.99999994)
    assert np.all(p[1] == np.float32(0.99999994))
    assert p[2].tolist() == [0.0, np.float32(0.99999994), 0.0]
    # the largest unit position never reaches the next cell at any level
    for r in S.level_resolutions()[0]:
        assert np.floor(np.float32(0.99999994) * np.float32(r)) == r - 1


# ------------------------------------------------------------------ BF16
def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.normal(size=2000).astype(np.float32) * 10 ** rng.uniform(-20, 20, 2000).astype(np.float32),
                        np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 0.0, -0.0], np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(S.bf16_round(x), ref)
    assert S.bf16_round(np.float32(1 + 2 ** -8)) == 1.0          # tie to even
    assert S.bf16_round(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6


# ------------------------------------------------------------------ MLP
def random_model(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    mlp = np.concatenate([rng.normal(size=n) * scale * math.sqrt(2.0 / f)
                          for n, f in zip(S.W_SIZES, (S.IN, S.HID, S.HID))])
    table = np.zeros((S.LEVELS, S.TABLE, 2))
    table[:] = 0.0
    return rng, mlp


def test_forward_matches_explicit_loops():
    rng, mlp = random_model(0)
    X = rng.normal(size=(5, S.IN))
    _, _, Y = S.forward(X, mlp, exact=False)
    W1, W2, W3 = S.split_mlp(mlp)
    for s in range(5):
        h1 = [max(0.0, sum(X[s, i] * W1[i, j] for i in range(S.IN))) for j in range(S.HID)]
        h2 = [max(0.0, sum(h1[i] * W2[i, j] for i in range(S.HID))) for j in range(S.HID)]
        y = [sum(h2[i] * W3[i, j] for i in range(S.HID)) for j in range(S.OUT)]
        np.testing.assert_allclose(Y[s], y, rtol=1e-12, atol=1e-12)


def fd_setup(n=6):
    rng, mlp = random_model(4)
    enc = S.Encoder(np.zeros(3), np.full(3, 3.0))
    pos = rng.uniform(0, 3, size=(n, 3))
    d = unit_dirs(n, 5)
    q = np.concatenate([pos, d, d], axis=1).astype(np.float32)
    table = rng.normal(size=(S.LEVELS, S.TABLE, 2)) * 0.5
    targets = rng.uniform(0.1, 2.0, size=(n, 3))
    samples = np.concatenate([q.astype(np.float64), targets], axis=1)
    return enc, table, mlp, samples


def test_error_head_trains_only_its_column():
    enc, table, mlp, samples = fd_setup()
    _, mg0, tg0, _ = S.train_gradients(enc, table, mlp, samples, 0.0, exact=False)
    _, mg1, tg1, fixed = S.train_gradients(enc, table, mlp, samples, 0.1, exact=False)
    a, b, _ = S.W_SIZES
    w3 = (mg1 - mg0)[a + b:].reshape(S.HID, S.OUT)
    assert np.all(w3[:, [0, 1, 2] + list(range(4, 16))] == 0) and np.any(w3[:, 3] != 0)
    np.testing.assert_array_equal(mg1[:a + b], mg0[:a + b])
    np.testing.assert_array_equal(tg1, tg0)
    # the column-3 gradient is the finite difference of the error term
    h = 1e-6
    for j in range(0, S.HID, 7):
        i = a + b + j * S.OUT + 3
        mp, mm = mlp.copy(), mlp.copy()
        mp[i] += h
        mm[i] -= h
        fd = (S.loss_only(enc, table, mp, samples, 0.1, fixed) - S.loss_only(enc, table, mm, samples, 0.1, fixed)) / (2 * h)
        assert abs(fd - mg1[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8


def test_gradients_match_finite_differences():
    # radiance terms only (error weight 0): every parameter reaches the loss
    enc, table, mlp, samples = fd_setup()
    ew = 0.0
    loss, mg, tg, fixed = S.train_gradients(enc, table, mlp, samples, ew, exact=False)
    assert abs(loss - S.loss_only(enc, table, mlp, samples, ew, fixed)) < 1e-12
    rng = np.random.default_rng(9)
    h = 1e-6
    # MLP weights: a spread over all three matrices, including nonzero grads
    picks = list(rng.choice(S.W_SIZES[0], 12, replace=False))
    picks += list(S.W_SIZES[0] + rng.choice(S.W_SIZES[1], 12, replace=False))
    picks += list(S.W_SIZES[0] + S.W_SIZES[1] + rng.choice(S.W_SIZES[2], 12, replace=False))
    picks += list(np.argsort(-np.abs(mg))[:8])
    for i in picks:
        mp, mm = mlp.copy(), mlp.copy()
        mp[i] += h
        mm[i] -= h
        fd = (S.loss_only(enc, table, mp, samples, ew, fixed) - S.loss_only(enc, table, mm, samples, ew, fixed)) / (2 * h)
        assert abs(fd - mg[i]) <= 1e-6 * max(1.0, abs(mg[i])) + 1e-8, (i, fd, mg[i])
    # table entries touched by the samples (dense and hashed levels)
    touched = np.argwhere(np.abs(tg) > 0)
    for l, e, f in touched[rng.choice(len(touched), 24, replace=False)]:
        tp, tm = table.copy(), table.copy()
        tp[l, e, f] += h
        tm[l, e, f] -= h
        fd = (S.loss_only(enc, tp, mlp, samples, ew, fixed) - S.loss_only(enc, tm, mlp, samples, ew, fixed)) / (2 * h)
        assert abs(fd - tg[l, e, f]) <= 1e-6 * max(1.0, abs(tg[l, e, f])) + 1e-8, (l, e, f, fd, tg[l, e, f])
    assert {l for l, _, _ in touched} == set(range(16))


def test_nonfinite_targets_are_skipped():
    enc, table, mlp, samples = fd_setup(4)
    s2 = samples.copy()
    s2[1, 9:12] = np.nan
    l_all, mg_all, _, _ = S.train_gradients(enc, table, mlp, samples, 0.1, exact=False)
    l_skip, mg_skip, _, _ = S.train_gradients(enc, table, mlp, s2, 0.1, exact=False)
    assert np.isfinite(l_skip) and np.all(np.isfinite(mg_skip)) and l_skip < l_all


def test_exact_mode_tracks_f64():
    enc, table, mlp, samples = fd_setup(64)
    X64 = enc.encode(samples[:, :9].astype(np.float32), table, exact=False)
    Xbf = enc.encode(samples[:, :9].astype(np.float32), table, exact=True)
    assert np.all(S.bf16_round(Xbf.astype(np.float32)) == Xbf.astype(np.float32))
    np.testing.assert_allclose(Xbf, X64, rtol=2 ** -8, atol=1e-3)
    _, _, Y64 = S.forward(X64, mlp, exact=False)
    _, _, Ybf = S.forward(Xbf, mlp, exact=True)
    assert np.linalg.norm(Ybf - Y64) <= 0.03 * np.linalg.norm(Y64)


def test_adam_first_step_is_signed_learning_rate():
    g = np.array([3.0, -1e-3, 0.0, 1e5])
    p, m, v = S.adam(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, 0.01)
    np.testing.assert_allclose(p, [-0.01, 0.01, 0.0, -0.01], rtol=1e-9)
    # the second step with the same gradient is the same signed step
    p2, _, _ = S.adam(p, g, m, v, 2, 0.01)
    np.testing.assert_allclose(p2 - p, [-0.01, 0.01, 0.0, -0.01], rtol=1e-9)
