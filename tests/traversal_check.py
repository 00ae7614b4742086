"""Shared acceptance logic for traversal results (reference semantics).

Follows the reference's GPU-vs-brute-force acceptance
(src/traversal_tests.rs:1462-1497): the distance must be within
2e-4 * max(1, t) of the nearest f64 candidate and the reported
(triangle, instance) must be one of the candidates within that tolerance,
so exact ties are tolerated; any-hit blockers must be some candidate.
"""
import numpy as np

EMPTY = 0xFFFFFFFF
ANY_HIT, FIRST_HIT = 1, 2
OCCLUDED, ROOT, UNSUPPORTED = 0x100, 0x200, 4


def slot_to_ids(built, tris, insts):
    tri_ids = built.tri[tris, 3].view(np.uint32) if len(tris) else np.zeros(0, np.uint32)
    inst_ids = built.inst[insts, 28] if len(insts) else np.zeros(0, np.uint32)
    return tri_ids, inst_ids


def check_against_brute_force(built, rays, hits, expected_errors=0, max_failures=12):
    failures = []
    counts = dict(closest=0, shadow=0, first=0)
    for i, (ray, h) in enumerate(zip(rays, hits)):
        flags = int(ray.view(np.uint32)[7])
        ts, tri_s, inst_s = built.brute_force(ray)
        cand_tri, cand_inst = slot_to_ids(built, tri_s, inst_s)
        t = float(h[0:1].view(np.float32)[0])
        tol = 0.0002 * max(abs(t), 1.0)
        err = int(h[3])
        if flags & ANY_HIT:
            blocked = bool(err & OCCLUDED)
            ok = blocked == (len(ts) > 0) and not (err & UNSUPPORTED)
            if ok and blocked:
                counts["shadow"] += 1
                ok = any(ct == h[1] and ci == h[2] and abs(t - c) <= tol
                         for c, ct, ci in zip(ts, cand_tri, cand_inst))
        else:
            hit = h[1] != EMPTY
            ok = (err & ~expected_errors) == 0 and (not hit or err == expected_errors)
            if not hit:
                ok = ok and len(ts) == 0
            elif flags & FIRST_HIT:
                counts["first"] += 1
                ok = ok and any(ct == h[1] and ci == h[2] and abs(t - c) <= tol
                                for c, ct, ci in zip(ts, cand_tri, cand_inst))
            elif len(ts):
                counts["closest"] += 1
                t0 = ts[0]
                ok = ok and abs(t - t0) <= tol and any(
                    abs(c - t0) <= tol and ct == h[1] and ci == h[2]
                    for c, ct, ci in zip(ts, cand_tri, cand_inst))
            else:
                ok = False
        if not ok and len(failures) < max_failures:
            failures.append(f"ray {i} {ray.tolist()}: hit {h.tolist()} t={t} brute {list(zip(ts[:3], cand_tri[:3], cand_inst[:3]))}")
    return failures, counts


def forced_deep(built, blas: bool):
    """The reference's forced 64-deep chains (src/traversal_tests.rs:1521-1593):
    every internal node's children share the root box, siblings are leaves over
    the whole primitive range, so a depth-first walk overflows the 49 stack."""
    import copy
    deep = copy.copy(built)
    internals = 64
    count = internals * 2 + 1

    def chain(nodes, mn, mx, first, n):
        root = len(nodes)
        for i in range(count):
            internal = i == 0 or (i % 2 == 1 and i < internals * 2 - 1)
            w = np.zeros(8, np.uint32)
            w[0:3] = mn.view(np.uint32)
            w[4:7] = mx.view(np.uint32)
            w[3] = root + (1 if i == 0 else i + 2) if internal else first
            w[7] = 0 if internal else n
            nodes.append(w)
        return root

    if blas:
        nodes, roots = [], {}
        inst = built.inst.copy()
        for k in range(inst.shape[0]):
            first, n, root = int(inst[k, 26]), int(inst[k, 27]), int(inst[k, 25])
            b = built.blas[root]
            mn, mx = b[0:3].copy(), b[4:7].copy()
            if first not in roots:
                roots[first] = chain(nodes, mn.view(np.float32), mx.view(np.float32), first, n)
            inst[k, 25] = roots[first]
        blas_nodes = np.asarray(nodes, np.uint32)
        from oracle.oracle import BuiltScene
        return BuiltScene(blas_nodes, built.tlas.copy(), built.tri.copy(), inst)
    nodes = []
    b = built.tlas[0]
    chain(nodes, b[0:3].view(np.float32), b[4:7].view(np.float32), 0, built.inst.shape[0])
    from oracle.oracle import BuiltScene
    return BuiltScene(built.blas.copy(), np.asarray(nodes, np.uint32), built.tri.copy(), built.inst.copy())


def first_hit_copies(rays):
    out = rays.copy()
    f = out.view(np.uint32)
    f[:, 7] = np.where(f[:, 7] == ANY_HIT, FIRST_HIT, f[:, 7])
    return out
