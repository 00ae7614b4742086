"""The reference's own CPU model of the split resolve and its fixtures.

The reference checks its GPU resolve against CPU helpers written in its
test module (src/world_resolve_tests.rs): a probe's 36-texel octahedral
irradiance table built from a radiance, a 32-bit mask of observed fixture
directions, a support and an angular pattern (reference_probe_table,
:77-127), the bilinear angular lookup (reference_angular_lookup, :129-167),
the trilinear diffuse lookup at one LOD over probes keyed by lattice
coordinates (reference_diffuse_at_lod, :169-235) and the LOD pair
(reference_split_diffuse_source / reference_split_diffuse, :237-309), with
the probe placement helper insert_probes (:433-480) and the fixture cases of
gpu_world_resolve_normalizes_sparse_support_and_isolates_lods (:749-980).
They are restated here in numpy float64 (background source only: the dense
path carries one radiance source), together with the placement of keyed
fixture probes into the dense LOD grids the HIP resolve reads.
"""
from __future__ import annotations

import math

import numpy as np

ALL = 0xFFFFFFFF


def fixture_direction(index: int) -> np.ndarray:
    """bin_direction of the test module (:31-38): 8 azimuths x 4 heights."""
    x, y = index % 8, index // 8
    phi = (x + 0.5) * 2.0 * math.pi / 8.0
    v = (y + 0.5) * 0.5 - 1.0
    r = math.sqrt(max(1.0 - v * v, 0.0))
    return np.array([r * math.cos(phi), v, r * math.sin(phi)])


def decode_octahedral(u: float, v: float) -> np.ndarray:
    px, py = u * 2.0 - 1.0, v * 2.0 - 1.0
    n = np.array([px, py, 1.0 - abs(px) - abs(py)])
    fold = min(max(-n[2], 0.0), 1.0)
    n[0] += -fold if n[0] >= 0.0 else fold
    n[1] += -fold if n[1] >= 0.0 else fold
    return n / np.linalg.norm(n)


def encode_octahedral(n) -> tuple:
    n = np.asarray(n, np.float64)
    n = n / np.linalg.norm(n)
    n = n / (abs(n[0]) + abs(n[1]) + abs(n[2]))
    px, py = n[0], n[1]
    if n[2] < 0.0:
        px, py = ((1.0 - abs(py)) * (1.0 if px >= 0.0 else -1.0),
                  (1.0 - abs(px)) * (1.0 if py >= 0.0 else -1.0))
    return px * 0.5 + 0.5, py * 0.5 + 0.5


def wrap_texel(x: int, y: int) -> tuple:
    if x < 0:
        x, y = -x - 1, 5 - y
    elif x >= 6:
        x, y = 11 - x, 5 - y
    if y < 0:
        x, y = 5 - x, -y - 1
    elif y >= 6:
        x, y = 5 - x, 11 - y
    return x, y


def probe_table(radiance, directions: int, support: float, pattern=None) -> np.ndarray:
    """(36, 4) table: per texel the cosine-weighted mean radiance over the
    observed fixture directions scaled to the nominal cosine sum x 4/32, and
    the observed / nominal cosine ratio as support."""
    pattern = np.ones(32) if pattern is None else np.asarray(pattern, np.float64)
    dirs = np.array([fixture_direction(d) for d in range(32)])
    s = min(max(float(support), 0.0), 1.0)
    out = np.zeros((36, 4))
    for t in range(36):
        n = decode_octahedral((t % 6) / 6.0 + 0.5 / 6.0, (t // 6) / 6.0 + 0.5 / 6.0)
        cos = np.maximum(dirs @ n, 0.0)
        nominal = cos.sum()
        mask = np.array([(directions >> d) & 1 for d in range(32)], bool)
        w = cos * s * mask
        observed = w.sum()
        if observed <= 0.0 or nominal <= 0.0:
            continue
        scale = nominal * (4.0 / 32.0) / observed
        rsum = np.asarray(radiance, np.float64) * (pattern * w).sum()
        out[t, :3] = rsum * scale
        out[t, 3] = min(max(observed / nominal, 0.0), 1.0)
    return out


def angular_lookup(table: np.ndarray, normal) -> tuple:
    u, v = encode_octahedral(normal)
    gx, gy = u * 6.0 - 0.5, v * 6.0 - 0.5
    bx, by = math.floor(gx), math.floor(gy)
    fx, fy = gx - bx, gy - by
    s = np.zeros(3)
    observed = 0.0
    for corner in range(4):
        cx, cy = corner & 1, corner >> 1
        bil = (fx if cx else 1.0 - fx) * (fy if cy else 1.0 - fy)
        x, y = wrap_texel(bx + cx, by + cy)
        smp = table[y * 6 + x]
        w = bil * min(max(smp[3], 0.0), 1.0)
        s += smp[:3] * w
        observed += w
    if observed <= 0.0:
        return np.zeros(3), 0.0
    return s / observed, min(max(observed, 0.0), 1.0)


def diffuse_at_lod(probes: dict, position, normal, lod: int) -> tuple:
    """probes: {(kx, ky, kz, lod): table}; lattice spacing 2^lod."""
    sp = 2.0 ** lod
    g = np.asarray(position, np.float64) / sp - 0.5
    base = np.floor(g).astype(int)
    fr = g - np.floor(g)
    s = np.zeros(3)
    observed = potential = 0.0
    for corner in range(8):
        bit = (corner & 1, (corner >> 1) & 1, (corner >> 2) & 1)
        spatial = np.prod([fr[a] if bit[a] else 1.0 - fr[a] for a in range(3)])
        if not spatial > 0.0:
            continue
        potential += spatial
        table = probes.get((base[0] + bit[0], base[1] + bit[1], base[2] + bit[2], lod))
        if table is None:
            continue
        irr, cov = angular_lookup(table, normal)
        if cov <= 0.0:
            continue
        s += irr * spatial * cov
        observed += spatial * cov
    if observed <= 0.0 or potential <= 0.0:
        return np.zeros(3), 0.0
    return s / observed, min(max(observed / potential, 0.0), 1.0)


def split_diffuse(probes: dict, position, normal, perspective: bool) -> tuple:
    """LOD pair of the fixture camera (LOD scale max|p| perspective, 1
    orthographic; LODs 0-4)."""
    scale = float(np.max(np.abs(position))) if perspective else 1.0
    low = int(min(max(math.floor(math.log2(max(scale, 1.0))), 0), 4))
    high, blend = low, 0.0
    if low < 4:
        nxt = 2.0 ** (low + 1)
        start = 0.9 * nxt
        if start <= scale < nxt:
            high = low + 1
            blend = min(max((scale - start) / (nxt - start), 0.0), 1.0)
    lv, lc = diffuse_at_lod(probes, position, normal, low)
    lw = lc * (1.0 - blend)
    s, total = lv * lw, lw
    if high != low:
        hv, hc = diffuse_at_lod(probes, position, normal, high)
        s = s + hv * hc * blend
        total += hc * blend
    if total <= 0.0:
        return np.zeros(3), 0.0
    return s / total, min(max(total, 0.0), 1.0)


def expected_alpha(coverage: float) -> float:
    return 1.0 + 0.25 * coverage if coverage > 0.0 else -1.0


def insert_probes(probes: dict, position, lod: int, radiance, directions: int, corners: int,
                  support: float = 1.0, pattern=None):
    """The 8 lattice probes around `position` at `lod` selected by the
    `corners` bit mask, each with the table of (radiance, directions,
    support, pattern); an existing probe is replaced."""
    sp = 2.0 ** lod
    base = np.floor(np.asarray(position, np.float64) / sp - 0.5).astype(int)
    for c in range(8):
        if not corners & (1 << c):
            continue
        key = (base[0] + (c & 1), base[1] + ((c >> 1) & 1), base[2] + ((c >> 2) & 1), lod)
        probes[key] = probe_table(radiance, directions, support, pattern)


def f32_tables(probes: dict) -> dict:
    """The tables as the HIP resolve reads them (float32)."""
    return {k: v.astype(np.float32).astype(np.float64) for k, v in probes.items()}


# dense LOD grids holding the fixtures: origin -4 (a multiple of every LOD's
# spacing up to LOD 2), LOD-0 dims 28 x 8 x 8 (x in [-4, 24], y, z in [-4, 4])
GRID_ORIGIN = (-4.0, -4.0, -4.0)
GRID_DIMS0 = (28, 8, 8)


def dense_tables(probes: dict, lods: int = 3) -> list:
    """(P_k, 36, 4) float32 packed tables of LOD k's dense grid: fixture
    probe (kx, ky, kz) sits at index k - origin / 2^lod; absent probes have
    zero support (a missing probe)."""
    out = []
    for lod in range(lods):
        sp = 2 ** lod
        dims = [-(-d // sp) for d in GRID_DIMS0]
        t = np.zeros((dims[2], dims[1], dims[0], 36, 4), np.float32)
        off = [int(-o // sp) for o in GRID_ORIGIN]
        for (kx, ky, kz, l), table in probes.items():
            if l != lod:
                continue
            ix, iy, iz = kx + off[0], ky + off[1], kz + off[2]
            assert 0 <= ix < dims[0] and 0 <= iy < dims[1] and 0 <= iz < dims[2], (kx, ky, kz, l)
            t[iz, iy, ix] = table.astype(np.float32)
        out.append(t.reshape(-1, 36, 4))
    return out


def grid_dims(lod: int) -> tuple:
    sp = 2 ** lod
    return tuple(-(-d // sp) for d in GRID_DIMS0)


def perspective_view_terms() -> dict:
    """perspective_infinite_reverse(pi/2, 1, 0.1) at the origin, viewport
    height 2, probe spacing 1, one target pixel (:521-527, :721-727): clip
    row 1 = (0, 1/tan(pi/4), 0), row 3 = (0, 0, -1)."""
    f = np.float32(1.0) / np.float32(math.tan(np.float32(math.pi / 2) * np.float32(0.5)))
    return dict(camera=(0.0, 0.0, 0.0), probe_spacing=1.0, clip_row1=(0.0, float(f), 0.0), viewport_height=2,
                clip_row3=(0.0, 0.0, -1.0), target_pixels=1)


def orthographic_view_terms() -> dict:
    """orthographic(-1, 1, -1, 1, 0.1, 100) at the origin (:714-719): clip
    row 1 = (0, 1, 0), row 3 = 0."""
    return dict(camera=(0.0, 0.0, 0.0), probe_spacing=1.0, clip_row1=(0.0, 1.0, 0.0), viewport_height=2,
                clip_row3=(0.0, 0.0, 0.0), target_pixels=1)


def fixture_cases() -> list:
    """(name, probes, perspective, receivers (n, 6)) of the reference's
    resolve fixture test (:749-980)."""
    Y = (0.0, 1.0, 0.0)
    cases = []
    pr = {}
    insert_probes(pr, (0, 0, 0), 0, [2.0] * 3, ALL, 0xFF)
    insert_probes(pr, (4, 0, 0), 0, [2.0] * 3, ALL, 0x33)
    insert_probes(pr, (8, 0, 0), 0, [2.0] * 3, (0x33 << 16) | (0x33 << 24), 0xFF)
    insert_probes(pr, (16, 0, 0), 0, [0.0] * 3, ALL, 0xFF)
    rec = [(x, 0.0, 0.0) + Y for x in (0.0, 4.0, 8.0, 16.0, 20.0)]
    cases.append(("sparse_support", pr, False, np.array(rec, np.float32)))
    pb = {}
    insert_probes(pb, (0, 0, 0), 0, [0.0] * 3, ALL, 0xFF, support=0.125)
    cases.append(("partial_black", pb, False, np.array([(0.0, 0.0, 0.0) + Y], np.float32)))
    tiny_pos = (1.9, 0.5009, 0.5009)
    tn = {}
    insert_probes(tn, tiny_pos, 0, [4.0] * 3, ALL, 1 << 6, support=1e-7)
    cases.append(("tiny_support", tn, True, np.array([tiny_pos + Y], np.float32)))
    overlap = (0.0, 0.0, -1.9)
    for lod, rad in ((0, 2.0), (1, 6.0)):
        f = {}
        insert_probes(f, overlap, lod, [rad] * 3, ALL, 0xFF)
        cases.append((f"overlap_lod{lod}", f, True, np.array([overlap + Y], np.float32)))
    both = {}
    insert_probes(both, overlap, 0, [2.0] * 3, ALL, 0xFF)
    insert_probes(both, overlap, 1, [6.0] * 3, ALL, 0xFF)
    cases.append(("overlap_both", both, True, np.array([overlap + Y], np.float32)))
    ang = {}
    insert_probes(ang, (0, 0, 0), 0, [2.0, 1.0, 3.0], 0xA55AC33C, 0xFF)
    normals = [np.array(Y), np.array([0.3, 0.2, 0.9]), np.array([0.001, 0.001, -1.0]),
               np.array([-0.001, -0.001, -1.0])]
    normals = [n / np.linalg.norm(n) for n in normals]
    cases.append(("angular", ang, False, np.array([(0.0, 0.0, 0.0) + tuple(n) for n in normals], np.float32)))
    zc = {}
    insert_probes(zc, (0, 0, 0), 0, [100.0] * 3, 0, 0xFF)
    cases.append(("zero_coverage", zc, False, np.array([(0.0, 0.0, 0.0) + Y], np.float32)))
    return cases


def close(actual: float, expected: float) -> bool:
    """The reference's tolerance (:729-734)."""
    return abs(actual - expected) <= max(abs(expected), 1.0) * 2e-4
