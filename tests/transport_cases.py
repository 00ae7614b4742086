"""Shared transport fixtures (inputs only; expectations live in the tests)."""
import numpy as np

from oracle import oracle as O


def child_direction(direction, level, child, base_theta=4):
    """testChildDirection (src/split_merge_tests.rs:631-636)."""
    theta = base_theta << level
    x, y = direction % (2 * theta), direction // (2 * theta)
    fx, fy = x * 2 + (child & 1), y * 2 + (child >> 1)
    return fx + fy * (4 * theta)


def four_cascade_case():
    """fourCascadeCases (src/split_merge_tests.rs:650-690) on a dense grid.

    Reference cells k at level L map to dense indices k + 8 / 2^L (grid origin
    at -8 level-0 cells), so level 3 cells -1..0 are indices 0..1, the level-2
    probe (0,0,0) is index 2, level 1 (1,1,1) is 5 and level 0 (3,3,3) is 11.
    Every other interval has no sample (beta = -1).
    """
    p = O.gi_params((-8.0, -8.0, -8.0), 1.0, (16, 16, 16), 4, base_phi=8, base_y=4)
    iv = np.zeros((O.total_intervals(p), 4), np.float32)
    iv[:, 3] = -1.0

    def put(level, cell, direction, radiance, beta):
        n, dims = O.level_dims(p, level)
        shift = 8 >> level
        x, y, z = (c + shift for c in cell)
        probe = (z * dims[1] + y) * dims[0] + x
        D = O.level_directions(p, level)
        g = O.level_offset(p, level) + probe * D + direction
        iv[g] = [radiance, radiance, radiance, beta]
        return g

    parents = [0, 1, 16, 17]
    level_two = [child_direction(d, 1, c) for d in parents for c in range(4)]
    for d2 in level_two:
        for c in range(4):
            d3 = child_direction(d2, 2, c)
            for corner in range(8):
                bit = (corner & 1, (corner >> 1) & 1, corner >> 2)
                put(3, tuple(-1 + b for b in bit), d3, 1.0, 0.0)
    g2 = [put(2, (0, 0, 0), d, 2.0, 0.5) for d in level_two]
    g1 = [put(1, (1, 1, 1), d, 4.0, 0.5) for d in parents]
    g0 = put(0, (3, 3, 3), 0, 8.0, 0.25)
    g3 = O.level_offset(p, 3) + 0 * O.level_directions(p, 3) + child_direction(level_two[0], 2, 0)
    return p, iv, dict(g3=g3, g2=g2[0], g1=g1[0], g0=g0)


def four_cascade_expected():
    """Closed forms of src/split_merge_tests.rs:1403-1419."""
    spatial = 0.75 ** 3
    s1 = 0.5 + 0.5 * spatial
    r1 = 4.0 + 0.5 * 2.5 * spatial
    s0 = 0.75 + 0.25 * spatial * s1
    r0 = 8.0 + 0.25 * spatial * r1
    return dict(g3=(1.0, 1.0), g2=(2.5, 1.0), g1=(r1 / s1, s1), g0=(r0 / s0, s0))


def cornell_params(**kw):
    """The small configuration: [0,2]^3 box, 16^3 probes at 0.125 m, 4 levels,
    16 level-0 directions (4 x 4), intervals [0,0.2],[0.2,0.6],[0.6,1.4],[1.4,3.0]."""
    args = dict(base_phi=4, base_y=4, interval_mode=1, interval_length=0.2, normal_bias=0.005,
                emissive_samples=16, seed=0, terminal_tail=1)
    args.update(kw)
    return O.gi_params((0.0, 0.0, 0.0), 0.125, (16, 16, 16), 4, **args)


def grid_receivers(n, lo, hi, seed=0):
    rng = np.random.default_rng(seed)
    rec = np.zeros((n, 6), np.float32)
    rec[:, :3] = lo + rng.random((n, 3)) * (np.asarray(hi) - np.asarray(lo))
    v = rng.normal(size=(n, 3))
    rec[:, 3:] = v / np.linalg.norm(v, axis=1, keepdims=True)
    rec[:8, 3:] = 0.0           # not eligible
    rec[8:12, 3:] = np.nan      # not eligible
    return rec
