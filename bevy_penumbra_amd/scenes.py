"""Seeded synthetic scenes and ray sets for the transport hot path.

Every generator is deterministic for a given seed and returns float32 arrays
in the layout the C ABI takes (non-indexed triangles, 9 floats each).

* ``colonnade``  -- the 262,144-triangle procedural colonnade (extruded boxes
  and arch segments, sizes log-uniform 0.05-4 m, 30 x 12 x 15 m, 8 emissive
  quads) used by the traversal microbenchmark and the split trace.
* ``traversal_fixture`` -- 24 instances of a cube, a panel and a bare mesh
  with the transforms, cull modes and mirrored scales of the reference's
  traversal GPU test scene (src/traversal_tests.rs:2120-2177).
* ``fixture_rays`` -- the reference's ray families for that scene
  (src/traversal_tests.rs:1040-1125): triangle interiors and shared
  diagonals from both faces at 0.01/1/100, 4,096 LCG rays (seed 7) and exact
  axis-parallel rays, each duplicated as an any-hit query.
* ``microbench_rays`` -- 4M rays: origins uniform in the scene box,
  directions uniform on the sphere, tmax log-uniform 0.1-20 m, first half
  closest-hit, second half any-hit.
* ``cornell`` -- the [0,2]^3 Cornell box of the small GI configuration.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

ANY_HIT = 1


@dataclass
class Scene:
    meshes: list            # list of (T, 9) float32
    mesh_materials: list    # list of (T,) uint32
    inst_inv: np.ndarray    # (I, 12) float32 local_from_world rows
    inst_fwd: np.ndarray    # (I, 12) float32 world_from_local rows
    inst_mesh: np.ndarray   # (I,) uint32
    inst_flags: np.ndarray  # (I,) uint32
    materials: np.ndarray   # (M, 8) float32: albedo.rgb, roughness, emission.rgb, pad
    bounds: tuple = field(default=((0.0, 0.0, 0.0), (1.0, 1.0, 1.0)))

    @property
    def triangle_count(self) -> int:
        return int(sum(m.shape[0] for m in self.meshes))


def _affine_rows(m4: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(m4[:3, :4].reshape(12)).astype(np.float32)


def single_instance(tris: np.ndarray, mats: np.ndarray, materials: np.ndarray, bounds) -> Scene:
    eye = np.eye(4)
    return Scene(
        meshes=[np.ascontiguousarray(tris, np.float32)],
        mesh_materials=[np.ascontiguousarray(mats, np.uint32)],
        inst_inv=_affine_rows(eye)[None, :].copy(),
        inst_fwd=_affine_rows(eye)[None, :].copy(),
        inst_mesh=np.zeros(1, np.uint32),
        inst_flags=np.zeros(1, np.uint32),
        materials=np.ascontiguousarray(materials, np.float32),
        bounds=bounds,
    )


# ---------------------------------------------------------------- primitives
_BOX_FACES = np.array(
    [  # quads over the 8 corners (bit0 x, bit1 y, bit2 z), outward winding
        [0, 2, 3, 1], [4, 5, 7, 6],  # -z, +z
        [0, 1, 5, 4], [2, 6, 7, 3],  # -y, +y
        [0, 4, 6, 2], [1, 3, 7, 5],  # -x, +x
    ],
    dtype=np.int64,
)


def hexahedron(corners: np.ndarray) -> np.ndarray:
    """12 triangles (12, 9) of a hexahedron given 8 corners (8, 3)."""
    tris = []
    for q in _BOX_FACES:
        a, b, c, d = corners[q]
        tris.append(np.concatenate([a, b, c]))
        tris.append(np.concatenate([a, c, d]))
    return np.asarray(tris)


def box_corners(center, half, yaw: float = 0.0) -> np.ndarray:
    sx, sy, sz = half
    k = np.arange(8)
    local = np.stack([np.where(k & 1, sx, -sx), np.where(k & 2, sy, -sy),
                      np.where(k & 4, sz, -sz)], axis=1).astype(np.float64)
    c, s = np.cos(yaw), np.sin(yaw)
    rot = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return local @ rot.T + np.asarray(center, np.float64)


def quad(p0, e1, e2) -> np.ndarray:
    p0, e1, e2 = (np.asarray(v, np.float64) for v in (p0, e1, e2))
    a, b, c, d = p0, p0 + e1, p0 + e1 + e2, p0 + e2
    return np.asarray([np.concatenate([a, b, c]), np.concatenate([a, c, d])])


def _clutter_boxes(rng, nbox: int, ext, smallest: float, largest: float) -> np.ndarray:
    """(12 * nbox, 9) triangles of boxes whose longest side is log-uniform in
    [smallest, largest] (the other two 10-100 % of it), random yaw, centres
    uniform where the box stays 0.05 m inside [0, ext]."""
    ext = np.asarray(ext, np.float64)
    longest = np.exp(rng.uniform(np.log(smallest), np.log(largest), size=nbox))
    aspect = rng.uniform(0.1, 1.0, size=(nbox, 3))
    aspect[np.arange(nbox), rng.integers(0, 3, nbox)] = 1.0
    size = longest[:, None] * aspect
    half = size / 2
    lo = half + 0.05
    hi = ext - half - 0.05
    hi = np.maximum(hi, lo)
    center = lo + rng.random((nbox, 3)) * (hi - lo)
    yaw = rng.uniform(0, 2 * np.pi, nbox)
    k = np.arange(8)
    sgn = np.stack([np.where(k & 1, 1.0, -1.0), np.where(k & 2, 1.0, -1.0),
                    np.where(k & 4, 1.0, -1.0)], axis=1)        # (8,3)
    local = sgn[None, :, :] * half[:, None, :]                  # (n,8,3)
    c, s = np.cos(yaw)[:, None], np.sin(yaw)[:, None]
    x = local[..., 0] * c + local[..., 2] * s
    z = -local[..., 0] * s + local[..., 2] * c
    corners = np.stack([x, local[..., 1], z], axis=-1) + center[:, None, :]
    faces = corners[:, _BOX_FACES]                              # (n,6,4,3)
    t1 = np.concatenate([faces[:, :, 0], faces[:, :, 1], faces[:, :, 2]], axis=-1)
    t2 = np.concatenate([faces[:, :, 0], faces[:, :, 2], faces[:, :, 3]], axis=-1)
    return np.stack([t1, t2], axis=2).reshape(nbox * 12, 9)


# ---------------------------------------------------------------- colonnade
def colonnade(seed: int = 0, triangles: int = 262_144):
    """The procedural colonnade scene.  Returns (Scene, emissive triangle ids)."""
    rng = np.random.default_rng(seed)
    ext = np.array([30.0, 12.0, 15.0])
    tris, mats = [], []
    materials = [
        [0.75, 0.75, 0.75, 1.0, 0, 0, 0, 0],   # 0 grey stone
        [0.70, 0.55, 0.40, 1.0, 0, 0, 0, 0],   # 1 sandstone
        [0.30, 0.35, 0.45, 1.0, 0, 0, 0, 0],   # 2 slate
        [0.8, 0.8, 0.8, 1.0, 17.0, 12.0, 4.0, 0],  # 3 emitter
    ]
    # floor and ceiling slabs
    for y, h in ((0.0, 0.2), (12.0 - 0.2, 0.2)):
        tris.append(hexahedron(box_corners((15.0, y + h / 2, 7.5), (15.0, h / 2, 7.5))))
        mats.append(np.full(12, 0, np.uint32))
    # columns: 3 rows along x
    rows_z = (3.0, 7.5, 12.0)
    cols_x = np.linspace(1.5, 28.5, 12)
    for z in rows_z:
        for x in cols_x:
            tris.append(hexahedron(box_corners((x, 5.0, z), (0.3, 4.8, 0.3))))
            mats.append(np.full(12, 1, np.uint32))
    # arch segments between adjacent columns of a row: bent hexahedra on a half ring
    segs = 16
    for z in rows_z:
        for i in range(len(cols_x) - 1):
            x0, x1 = cols_x[i] + 0.3, cols_x[i + 1] - 0.3
            cx, r_in = 0.5 * (x0 + x1), 0.5 * (x1 - x0)
            r_out = r_in + 0.35
            for k in range(segs):
                a0, a1 = np.pi * k / segs, np.pi * (k + 1) / segs
                pts = []
                for j in range(8):
                    ang = a1 if j & 1 else a0
                    rad = r_out if j & 2 else r_in
                    zz = z + (0.3 if j & 4 else -0.3)
                    pts.append([cx - rad * np.cos(ang), 9.8 + rad * np.sin(ang) * 0.6, zz])
                tris.append(hexahedron(np.asarray(pts)))
                mats.append(np.full(12, 1, np.uint32))
    # 8 emissive quads near the ceiling, facing down
    emit_first = sum(t.shape[0] for t in tris)
    for q in range(8):
        cx = 3.0 + 3.4 * q
        cz = 4.0 + 7.0 * (q % 2)
        sx, sz = rng.uniform(0.8, 1.6), rng.uniform(0.8, 1.6)
        p0 = (cx - sx / 2, 11.7, cz - sz / 2)
        tris.append(quad(p0, (sx, 0, 0), (0, 0, sz)))  # normal -y
        mats.append(np.full(2, 3, np.uint32))
    emit_ids = np.arange(emit_first, emit_first + 16, dtype=np.uint32)
    used = sum(t.shape[0] for t in tris)
    remaining = triangles - used
    assert remaining >= 0 and remaining % 12 == 0, (used, remaining)
    nbox = remaining // 12
    # clutter: extruded boxes whose longest side is log-uniform in 0.05-4 m
    # (the other two sides 10-100 % of it), random yaw, kept inside the box
    clutter = _clutter_boxes(rng, nbox, ext, 0.05, 4.0)
    tris.append(clutter)
    mats.append(np.repeat(rng.integers(0, 3, nbox).astype(np.uint32), 12))
    all_tris = np.concatenate(tris).astype(np.float32)
    all_mats = np.concatenate(mats).astype(np.uint32)
    assert all_tris.shape[0] == triangles
    bounds = ((0.0, 0.0, 0.0), tuple(ext))
    return single_instance(all_tris, all_mats, np.asarray(materials, np.float32), bounds), emit_ids


def microbench_rays(bounds, n: int = 4 * 1024 * 1024, seed: int = 1) -> np.ndarray:
    """(n, 8) float32 rays; rows [n/2, n) are any-hit queries."""
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(bounds[0]), np.asarray(bounds[1])
    o = lo + rng.random((n, 3)) * (hi - lo)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    tmax = np.exp(rng.uniform(np.log(0.1), np.log(20.0), n))
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = tmax
    rays[:, 4:7] = v
    flags = rays.view(np.uint32)
    flags[:, 7] = 0
    flags[n // 2:, 7] = ANY_HIT
    return rays


# ---------------------------------------------------------------- fixture scene
def _euler_xyz(a, b, c):
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    rx = np.array([[1, 0, 0], [0, ca, -sa], [0, sa, ca]])
    ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    rz = np.array([[cc, -sc, 0], [sc, cc, 0], [0, 0, 1]])
    return rx @ ry @ rz


def traversal_fixture() -> Scene:
    """Cube / panel / bare meshes under 24 transforms (cull modes none, back, front)."""
    cube = hexahedron(box_corners((0, 0, 0), (0.5, 0.5, 0.5)))
    # 2 x 2 rectangle in the XY plane, corners (+,+) (-,+) (-,-) (+,-), two
    # triangles sharing corner 0 and the diagonal; the bare mesh is the same
    # rectangle without attributes (src/traversal_tests.rs:2082-2089)
    rect = np.asarray([[1, 1, 0], [-1, 1, 0], [-1, -1, 0], [1, -1, 0]], np.float64)
    panel = np.asarray([np.concatenate(rect[[0, 1, 2]]), np.concatenate(rect[[0, 2, 3]])])
    bare = panel.copy()
    meshes = [cube.astype(np.float32), panel.astype(np.float32), bare.astype(np.float32)]
    mats = [np.zeros(m.shape[0], np.uint32) for m in meshes]
    inv, fwd, mesh, flags = [], [], [], []
    cull_flags = [0, 1, 2]  # None, Back, Front
    for i in range(24):
        m = 2 if i == 23 else (0 if i % 2 == 0 else 1)
        if i == 23:
            t = np.array([-16.0, 0.0, -20.0]); r = np.eye(3); s = np.ones(3)
        else:
            t = np.array([(i % 6) * 4.0 - 10.0, (i // 6) * 5.0 - 7.5, 0.0])
            r = np.eye(3) if i == 0 else _euler_xyz(i * 0.03, i * 0.07, 0.2)
            s = np.array([-0.001, 1000.0, 1.0]) if i == 0 else \
                np.array([-1.2 if i % 2 == 0 else 1.2, 0.6, 2.0])
        mtx = np.eye(4)
        mtx[:3, :3] = r @ np.diag(s)
        mtx[:3, 3] = t
        # round the forward transform to f32 first, then invert it in f64
        m32 = mtx.astype(np.float32).astype(np.float64)
        fwd.append(_affine_rows(m32))
        inv.append(_affine_rows(np.linalg.inv(m32)))
        mesh.append(m)
        f = cull_flags[i % 3]
        if np.linalg.det(m32[:3, :3]) < 0:
            f |= 4
        flags.append(f)
    materials = np.asarray([[0.8, 0.6, 0.4, 1.0, 0.1, 0.2, 0.3, 0.0]], np.float32)
    return Scene(meshes, mats, np.asarray(inv), np.asarray(fwd), np.asarray(mesh, np.uint32),
                 np.asarray(flags, np.uint32), materials, bounds=((-60, -60, -60), (60, 60, 60)))


def _world_tris(scene: Scene, i: int) -> np.ndarray:
    m = scene.meshes[int(scene.inst_mesh[i])].astype(np.float32).reshape(-1, 3, 3)
    f = scene.inst_fwd[i].reshape(3, 4).astype(np.float32)
    return (m @ f[:, :3].T + f[:, 3]).astype(np.float32)


def fixture_rays(scene: Scene) -> np.ndarray:
    """The reference's ray families for a scene (closest-hit and any-hit copies)."""
    out = []
    f32 = np.float32
    for i in range(scene.inst_mesh.shape[0]):
        local = scene.meshes[int(scene.inst_mesh[i])].reshape(-1, 3, 3)[:12]
        world = _world_tris(scene, i)[:12]
        f = scene.inst_fwd[i].reshape(3, 4).astype(np.float64)
        identity = np.allclose(f[:, :3], np.eye(3))
        for lt, wt in zip(local, world):
            a, b, c = wt.astype(f32)
            n = np.cross(b - a, c - a).astype(f32)
            nl = np.linalg.norm(n)
            if not nl > 0:
                continue
            n = (n / nl).astype(f32)
            center = ((a + b + c) / f32(3)).astype(f32)
            edges = [(lt[0], lt[1]), (lt[1], lt[2]), (lt[2], lt[0])]
            s0, e0 = max(edges, key=lambda e: float(np.sum((e[1] - e[0]) ** 2)))
            mid = ((s0 + e0) * 0.5).astype(np.float64)
            diag = (f[:, :3] @ mid + f[:, 3]).astype(f32)
            targets = [center, diag, a] if identity else [center, diag]
            for tgt in targets:
                for sign in (-1.0, 1.0):
                    for dist in (0.01, 1.0, 100.0):
                        o = (tgt + n * f32(sign * dist)).astype(f32)
                        d = (tgt - o).astype(f32)
                        dl = np.linalg.norm(d)
                        if not (dl > 0 and np.isfinite(dl)):
                            continue
                        d = (d / dl).astype(f32)
                        out.append([*o, dist * 1.1, *d, 0.0])
    state = np.uint32(7)
    vals = []
    for _ in range(4096 * 6):
        state = np.uint32((int(state) * 1664525 + 1013904223) & 0xFFFFFFFF)
        vals.append(np.float32(int(state) >> 8) / np.float32(16777216.0))
    vals = np.asarray(vals, np.float32).reshape(4096, 6)
    for r in vals:
        o = r[:3] * f32(100.0) - f32(50.0)
        d = r[3:] * f32(2.0) - f32(1.0)
        d = (d / np.linalg.norm(d)).astype(f32)
        out.append([*o, 200.0, *d, 0.0])
    for i in range(scene.inst_mesh.shape[0]):
        w = _world_tris(scene, i).reshape(-1, 3)
        center = ((w.min(0) + w.max(0)) * 0.5).astype(f32)
        for d in np.concatenate([np.eye(3), -np.eye(3)]).astype(f32):
            o = center - d * f32(100.0)
            out.append([*o, 200.0, *d, 0.0])
    rays = np.asarray(out, np.float32)
    shadows = rays.copy()
    shadows.view(np.uint32)[:, 7] = ANY_HIT
    return np.ascontiguousarray(np.concatenate([rays, shadows]))


# ---------------------------------------------------------------- Cornell box
def cornell():
    """[0,2]^3 Cornell box: 5 walls (10 tris), short block 0.6 m (yaw 17 deg),
    tall block 0.6 x 1.2 x 0.6 (yaw -20 deg), 0.5 x 0.5 m ceiling emitter with
    radiance (17, 12, 4).  Red left wall, green right wall, grey elsewhere.
    Returns (Scene, emissive triangle ids)."""
    materials = np.asarray([
        [0.75, 0.75, 0.75, 1.0, 0, 0, 0, 0],     # 0 grey
        [0.63, 0.065, 0.05, 1.0, 0, 0, 0, 0],    # 1 red
        [0.14, 0.45, 0.091, 1.0, 0, 0, 0, 0],    # 2 green
        [0.75, 0.75, 0.75, 1.0, 17.0, 12.0, 4.0, 0],  # 3 emitter
    ], np.float32)
    tris, mats = [], []
    def add(t, m):
        tris.append(np.asarray(t, np.float64)); mats.append(np.full(len(t), m, np.uint32))
    add(quad((0, 0, 0), (0, 0, 2), (2, 0, 0)), 0)      # floor, normal +y
    add(quad((0, 2, 0), (2, 0, 0), (0, 0, 2)), 0)      # ceiling, normal -y
    add(quad((0, 0, 0), (2, 0, 0), (0, 2, 0)), 0)      # back wall z=0, normal +z
    add(quad((0, 0, 0), (0, 2, 0), (0, 0, 2)), 1)      # left x=0 red, normal +x
    add(quad((2, 0, 0), (0, 0, 2), (0, 2, 0)), 2)      # right x=2 green, normal -x
    add(hexahedron(box_corners((1.3, 0.3, 1.3), (0.3, 0.3, 0.3), np.radians(17.0))), 0)
    add(hexahedron(box_corners((0.65, 0.6, 0.7), (0.3, 0.6, 0.3), np.radians(-20.0))), 0)
    first = sum(t.shape[0] for t in tris)
    add(quad((0.75, 1.999, 0.75), (0.5, 0, 0), (0, 0, 0.5)), 3)  # emitter, normal -y
    all_tris = np.concatenate(tris).astype(np.float32)
    all_mats = np.concatenate(mats)
    return (single_instance(all_tris, all_mats, materials, ((0, 0, 0), (2, 2, 2))),
            np.arange(first, first + 2, dtype=np.uint32))


# ---------------------------------------------------------------- dynamic frame
def _uv_sphere(stacks: int, slices: int) -> np.ndarray:
    """(2 * stacks * slices, 3, 3) triangles of a unit UV sphere, outward
    winding, band by band from the +y pole (the pole bands keep one
    zero-area triangle per quad so every band has 2 * slices triangles)."""
    th = np.linspace(0.0, np.pi, stacks + 1)
    ph = np.linspace(0.0, 2 * np.pi, slices + 1)
    T, P = np.meshgrid(th, ph, indexing="ij")
    v = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], axis=-1)
    a, b = v[:-1, :-1], v[1:, :-1]
    c, d = v[1:, 1:], v[:-1, 1:]
    t1 = np.stack([a, c, b], axis=2)
    t2 = np.stack([a, d, c], axis=2)
    return np.stack([t1, t2], axis=2).reshape(-1, 3, 3)


def dynamic(seed: int = 0, meshes: int = 256, instances: int = 256, stacks: int = 32, slices: int = 64,
            emissive_every: int = 1):
    """The dynamic large frame: ``instances`` rigid instances of ``meshes``
    procedural meshes of 2 * stacks * slices triangles (defaults: one mesh per
    instance, 256 x 4096 = 1,048,576 triangles) moving inside a 30 x 12 x 15 m
    volume.  Each mesh is a
    UV sphere with seeded radial relief, scaled to an ellipsoid of 0.3-1.6 m
    radii; the band below its top pole (2 * slices triangles) is emissive (on
    every emissive_every-th mesh).  Returns
    (Scene at frame 0, motion) where ``motion`` feeds ``dynamic_transforms``."""
    rng = np.random.default_rng(seed)
    base = _uv_sphere(stacks, slices)
    mesh_list, mat_list = [], []
    for m in range(meshes):
        k = rng.integers(1, 5, size=3)
        amp = rng.uniform(0.0, 0.15, size=3)
        ph = rng.uniform(0, 2 * np.pi, size=3)
        p = base.reshape(-1, 3)
        r = 1.0 + amp[0] * np.sin(k[0] * np.arctan2(p[:, 2], p[:, 0]) + ph[0]) \
            + amp[1] * np.sin(k[1] * np.arccos(np.clip(p[:, 1], -1, 1)) + ph[1]) \
            + amp[2] * np.sin(k[2] * (p[:, 0] + p[:, 1]) + ph[2])
        radii = rng.uniform(0.3, 1.6, size=3)
        tri = (p * r[:, None] * radii[None, :]).reshape(-1, 9).astype(np.float32)
        mesh_list.append(tri)
        mats = np.full(tri.shape[0], m % 3, np.uint32)
        if m % emissive_every == 0:
            mats[2 * slices: 4 * slices] = 3
        mat_list.append(mats)
    materials = np.asarray([
        [0.75, 0.75, 0.75, 1.0, 0, 0, 0, 0],
        [0.70, 0.55, 0.40, 1.0, 0, 0, 0, 0],
        [0.30, 0.35, 0.45, 1.0, 0, 0, 0, 0],
        [0.8, 0.8, 0.8, 1.0, 6.0, 5.0, 3.0, 0],
    ], np.float32)
    ext = np.array([30.0, 12.0, 15.0])
    axis = rng.normal(size=(instances, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    motion = dict(
        centre=rng.uniform([2.0, 1.5, 2.0], ext - [2.0, 1.5, 2.0], size=(instances, 3)),
        sway=rng.uniform(0.1, 1.0, size=(instances, 3)),
        freq=rng.uniform(0.01, 0.05, size=(instances, 3)),
        phase=rng.uniform(0, 2 * np.pi, size=(instances, 3)),
        axis=axis,
        spin=rng.uniform(-0.05, 0.05, size=instances),
        angle0=rng.uniform(0, 2 * np.pi, size=instances),
        mesh=(np.arange(instances) % meshes).astype(np.uint32),
    )
    inv, fwd = dynamic_transforms(motion, 0)
    scene = Scene(meshes=mesh_list, mesh_materials=mat_list, inst_inv=inv, inst_fwd=fwd,
                  inst_mesh=motion["mesh"].copy(), inst_flags=np.zeros(instances, np.uint32),
                  materials=materials, bounds=((0.0, 0.0, 0.0), tuple(ext)))
    return scene, motion


def dynamic_transforms(motion: dict, frame: int):
    """Seeded rigid motion of every instance at a frame: rotation about a
    fixed axis at a fixed rate, translation swaying around a centre.  Returns
    (local_from_world, world_from_local) as (I, 12) float32 row-major 3x4."""
    f = float(frame)
    t = motion["centre"] + motion["sway"] * np.sin(motion["freq"] * f + motion["phase"])
    ang = motion["angle0"] + motion["spin"] * f
    a = motion["axis"]
    c, s = np.cos(ang)[:, None, None], np.sin(ang)[:, None, None]
    K = np.zeros((a.shape[0], 3, 3))
    K[:, 0, 1], K[:, 0, 2] = -a[:, 2], a[:, 1]
    K[:, 1, 0], K[:, 1, 2] = a[:, 2], -a[:, 0]
    K[:, 2, 0], K[:, 2, 1] = -a[:, 1], a[:, 0]
    R = np.eye(3)[None] * c + s * K + (1 - c) * a[:, :, None] * a[:, None, :]
    fwd = np.concatenate([R, t[:, :, None]], axis=2)
    Rt = np.transpose(R, (0, 2, 1))
    inv = np.concatenate([Rt, -(Rt @ t[:, :, None])], axis=2)
    return (np.ascontiguousarray(inv.reshape(-1, 12), np.float32),
            np.ascontiguousarray(fwd.reshape(-1, 12), np.float32))


def camera_rays(origin, target, width: int, height: int, fov_y_deg: float = 60.0) -> np.ndarray:
    """Pinhole camera basis for ``lux_camera_receivers``: 12 floats = origin,
    forward (unit), right and up scaled to the half extents of the image
    plane at unit distance."""
    o = np.asarray(origin, np.float64)
    fwd = np.asarray(target, np.float64) - o
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    h = np.tan(np.radians(fov_y_deg) / 2)
    w = h * width / height
    return np.concatenate([o, fwd, right * w, up * h]).astype(np.float32)


# ---------------------------------------------------------------- batched rooms
def clutter_room(seed: int, triangles: int, size: float = 8.0):
    """A closed cubic room of side `size` m (floor, ceiling and four walls as
    0.1 m slabs), 2-4 emissive ceiling quads and boxes of log-uniform size
    0.05-1.5 m filling it to `triangles` (rounded down to whole boxes, at
    least 84 triangles).  Returns (Scene, emissive triangle ids)."""
    rng = np.random.default_rng(seed)
    L = float(size)
    materials = [
        [0.75, 0.75, 0.75, 1.0, 0, 0, 0, 0],
        [0.63, 0.20, 0.15, 1.0, 0, 0, 0, 0],
        [0.20, 0.45, 0.30, 1.0, 0, 0, 0, 0],
        [0.8, 0.8, 0.8, 1.0, 17.0, 12.0, 4.0, 0],
    ]
    tris, mats = [], []
    t = 0.05
    for c, h in (((L / 2, t, L / 2), (L / 2, t, L / 2)), ((L / 2, L - t, L / 2), (L / 2, t, L / 2)),
                 ((t, L / 2, L / 2), (t, L / 2, L / 2)), ((L - t, L / 2, L / 2), (t, L / 2, L / 2)),
                 ((L / 2, L / 2, t), (L / 2, L / 2, t)), ((L / 2, L / 2, L - t), (L / 2, L / 2, t))):
        tris.append(hexahedron(box_corners(c, h)))
        mats.append(np.full(12, len(mats) % 3, np.uint32))
    emit_first = sum(x.shape[0] for x in tris)
    n_emit = int(rng.integers(2, 5))
    for _ in range(n_emit):
        sx, sz = rng.uniform(0.1, 0.25, 2) * L
        x0, z0 = rng.uniform(0.1 * L, 0.9 * L - sx), rng.uniform(0.1 * L, 0.9 * L - sz)
        tris.append(quad((x0, L - 2 * t - 1e-3, z0), (sx, 0, 0), (0, 0, sz)))  # normal -y
        mats.append(np.full(2, 3, np.uint32))
    emit_ids = np.arange(emit_first, emit_first + 2 * n_emit, dtype=np.uint32)
    used = sum(x.shape[0] for x in tris)
    nbox = max(1, (int(triangles) - used) // 12)
    tris.append(_clutter_boxes(rng, nbox, (L, 0.85 * L, L), 0.05, 1.5))
    mats.append(np.repeat(rng.integers(0, 3, nbox).astype(np.uint32), 12))
    all_tris = np.concatenate(tris).astype(np.float32)
    all_mats = np.concatenate(mats).astype(np.uint32)
    return (single_instance(all_tris, all_mats, np.asarray(materials, np.float32), ((0.0, 0.0, 0.0), (L, L, L))),
            emit_ids)


def batch_set(seed: int = 0, count: int = 64, triangles=(1_000, 200_000), grid=(8, 48), levels=(3, 5)):
    """Seeded specifications of `count` uneven scenes: triangle counts and
    cubic probe-grid sides log-uniform in the given ranges, cascade counts
    uniform in `levels`, room sides 4-12 m.  Each entry: seed, triangles,
    size (m), dims, spacing (m), levels."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        tri = int(np.exp(rng.uniform(np.log(triangles[0]), np.log(triangles[1]))))
        d = int(round(np.exp(rng.uniform(np.log(grid[0]), np.log(grid[1])))))
        lv = int(rng.integers(levels[0], levels[1] + 1))
        size = float(rng.uniform(4.0, 12.0))
        out.append(dict(seed=int(seed) * 1000 + i, triangles=tri, size=size, dims=(d, d, d),
                        spacing=size / d, levels=lv))
    return out
