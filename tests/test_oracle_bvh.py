"""Pins the oracle's binned-SAH builder, BVH2 layout, BLAS/TLAS packing and refit.

Golden cases follow the reference's tests/unit/binnedSahBvh.test.ts:71-260;
the root split is also recomputed independently in numpy (SAH over all bin
positions) so that a wrong sweep index or cost term fails.
"""
import math

import numpy as np
import pytest


def park_miller_boxes(count):
    seed = 1

    def rnd():
        nonlocal seed
        seed = (seed * 16807) % 2147483647
        return seed / 2147483647

    out = np.zeros((count, 6), np.float32)
    for i in range(count):
        x, y, z, s = rnd() * 100, rnd() * 100, rnd() * 100, rnd()
        out[i] = [x, y, z, x + s, y + s, z + s]
    return out


@pytest.mark.parametrize("bins", [8, 16, 32, 64])
def test_every_primitive_once_children_inside(orc, bins):
    count = 1000
    aabbs = park_miller_boxes(count)
    bvh = orc.build_binned_sah(aabbs, bins)
    assert sorted(bvh["prim_indices"].tolist()) == list(range(count))
    leaf_prims = 0
    b = bvh["bounds"]
    for node in range(bvh["node_count"]):
        if bvh["prim_count"][node]:
            leaf_prims += int(bvh["prim_count"][node])
            continue
        left = int(bvh["left_first"][node])
        for child in (left, left + 1):
            assert np.all(b[child, :3] >= b[node, :3])
            assert np.all(b[child, 3:] <= b[node, 3:])
    assert leaf_prims == count
    assert bvh["node_count"] < 2 * count
    assert bvh["vec4_count"] == bvh["node_count"] * 2 + count
    assert bvh["order"][0] == 0


def test_bin_edge_centroid_partition(orc):
    f = np.float32
    lo, hi, c = -1000.0, float(f(1000.3)), float(f(0.14999441802501678))
    boxes = [lo, 0, 0, lo + 0.01, 1, 1, hi - 0.01, 0, 0, hi, 1, 1, c, 0, 0, c, 1, 1]
    for b in range(8):
        x = lo + ((b + 0.5) * (hi - lo)) / 8
        boxes += [x - 0.01, 0, 0, x + 0.01, 1, 1]
    aabbs = np.array(boxes, np.float32).reshape(-1, 6)
    bvh = orc.build_binned_sah(aabbs)

    def contains(node):
        if bvh["prim_count"][node]:
            s = int(bvh["left_first"][node])
            prims = bvh["prim_indices"][s:s + int(bvh["prim_count"][node])].tolist()
        else:
            l = int(bvh["left_first"][node])
            prims = contains(l) + contains(l + 1)
        for p in prims:
            assert np.all(aabbs[p, :3] >= bvh["bounds"][node, :3])
            assert np.all(aabbs[p, 3:] <= bvh["bounds"][node, 3:])
        return prims

    assert len(contains(0)) == aabbs.shape[0]


def test_coincident_bounds_single_leaf(orc):
    bvh = orc.build_binned_sah(np.tile(np.array([0, 0, 0, 1, 1, 1], np.float32), (5, 1)))
    assert bvh["node_count"] == 1 and bvh["prim_count"][0] == 5


def test_sah_stop_rule(orc):
    assert orc.build_binned_sah(np.array([[0, 0, 0, 1, 1, 1], [0, 0, 0, 1, 1, 1.001]], np.float32))["node_count"] == 1
    assert orc.build_binned_sah(np.array([[0, 0, 0, 1, 1, 1], [9, 0, 0, 10, 1, 1]], np.float32))["node_count"] == 3


def test_empty(orc):
    assert orc.build_binned_sah(np.zeros((0, 6), np.float32))["node_count"] == 0


def _root_split_numpy(aabbs, bins):
    """Independent numpy SAH over every bin position of the root."""
    a = aabbs.astype(np.float64)
    c = ((a[:, :3] + a[:, 3:]) * 0.5).astype(np.float32).astype(np.float64)
    lo, hi = a[:, :3].min(0), a[:, 3:].max(0)
    ext = hi - lo
    area = ext[0] * ext[1] + ext[1] * ext[2] + ext[2] * ext[0]
    best = (math.inf, None, None)
    for axis in range(3):
        if not ext[axis] > ext[axis] * 1e-20:
            continue
        scale = bins / ext[axis]
        b = np.clip(np.trunc((c[:, axis] - lo[axis]) * scale), 0, bins - 1).astype(int)
        for pos in range(bins - 1):
            L, R = a[b <= pos], a[b > pos]
            if len(L) == 0 or len(R) == 0:
                continue

            def half_area(s):
                e = s[:, 3:].max(0) - s[:, :3].min(0)
                return e[0] * e[1] + e[1] * e[2] + e[2] * e[0]

            cost = half_area(L) * len(L) + half_area(R) * len(R)
            if cost < best[0]:
                best = (cost, axis, pos, len(L))
    return 1 + best[0] / area, best[3]


@pytest.mark.parametrize("bins", [8, 32])
def test_root_split_matches_exhaustive_sah(orc, bins):
    rng = np.random.default_rng(11)
    cen = rng.random((3000, 3)) * np.array([10, 3, 1])
    size = rng.lognormal(-3, 0.5, (3000, 1))
    aabbs = np.concatenate([cen - size, cen + size], 1).astype(np.float32)
    bvh = orc.build_binned_sah(aabbs, bins)
    cost, left_count = _root_split_numpy(aabbs, bins)
    assert cost < 3000
    left = int(bvh["left_first"][0])
    assert int(bvh["prim_count"][left]) == left_count or int(bvh["prim_count"][left]) == 0
    # child sizes from the subtree leaf counts
    def leaves(node):
        if bvh["prim_count"][node]:
            return int(bvh["prim_count"][node])
        l = int(bvh["left_first"][node])
        return leaves(l) + leaves(l + 1)
    assert leaves(left) == left_count


def sphere(seg=32, rings=16):
    verts, idx = [], []
    for r in range(rings + 1):
        th = math.pi * r / rings
        for s in range(seg + 1):
            ph = 2 * math.pi * s / seg
            verts.append([math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)])
    for r in range(rings):
        for s in range(seg):
            a, b = r * (seg + 1) + s, (r + 1) * (seg + 1) + s
            idx += [a, b, a + 1, b, b + 1, a + 1]
    return np.array(verts, np.float32), np.array(idx, np.uint32)


def validate_blas(flat, positions):
    w = flat["words"]
    f = w.view(np.float32)
    i = w.view(np.int32)
    entries = []
    depth_max = [0]
    leaf_max = [0]

    def visit(index, depth):
        depth_max[0] = max(depth_max[0], depth)
        base = index * 4
        child = int(i[base + 3])
        exp = np.array([np.inf] * 3 + [-np.inf] * 3)
        if child >= 0:
            assert child > index + 2
            for bb in (visit(index + 2, depth + 1), visit(child, depth + 1)):
                exp[:3] = np.minimum(exp[:3], bb[:3])
                exp[3:] = np.maximum(exp[3:], bb[3:])
        else:
            leaf_max[0] = max(leaf_max[0], -child)
            for k in range(-child):
                e = i[base + 8 + k * 4: base + 12 + k * 4].tolist()
                entries.append(e)
                for j in range(3):
                    p = positions[e[j]]
                    exp[:3] = np.minimum(exp[:3], p)
                    exp[3:] = np.maximum(exp[3:], p)
        act = np.array([f[base], f[base + 1], f[base + 2], f[base + 4], f[base + 5], f[base + 6]])
        assert np.array_equal(act.astype(np.float32), exp.astype(np.float32))
        return act

    if w.size:
        visit(0, 1)
    assert flat["max_depth"] == depth_max[0]
    assert flat["max_leaf"] == leaf_max[0]
    return entries


def test_blas_tight_and_complete(orc):
    pos, idx = sphere(32, 16)
    flat = orc.build_blas(pos, idx)
    entries = validate_blas(flat, pos)
    exp = sorted(",".join(map(str, idx[t * 3:t * 3 + 3])) for t in range(idx.size // 3))
    assert sorted(",".join(map(str, e[:3])) for e in entries) == exp


def test_blas_skips_non_finite(orc):
    pos = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, np.nan, 0, 0, 3, 0, 0, 2, 1, 0], np.float32).reshape(-1, 3)
    flat = orc.build_blas(pos, np.arange(6, dtype=np.uint32))
    assert [e[:3] for e in validate_blas(flat, pos)] == [[0, 1, 2]]
    empty = orc.build_blas(np.array([np.nan, 0, 0, 1, 0, 0, 0, 1, 0], np.float32), np.arange(3, dtype=np.uint32))
    assert empty["words"].size == 0 and empty["max_depth"] == 0


@pytest.mark.parametrize("bins", [8, 32])
def test_tlas_records_once(orc, bins):
    n = 50
    bounds = np.array([[i * 3, 0, 0, i * 3 + 1, 1, 1] for i in range(n)], np.float32)
    tr = np.tile(np.eye(4).reshape(16), (n, 1))
    tr[:, 12] = np.arange(n)
    inst = dict(blas_index=np.arange(n) % 2, bounds=bounds, transform=tr, material_word=np.arange(n))
    h = orc.build_tlas(inst, blas_offsets=[0, 3], vertex_offsets=[0, 3], bins=bins, per_instance_bounds=False)
    t = h.get()
    ints = t["words"].view(np.int32)
    rec = t["instance_words"].view(np.int32)
    recf = t["instance_words"].view(np.float32)
    seen = []

    def visit(index):
        child = ints[index * 4 + 3]
        if child >= 0:
            visit(index + 2)
            visit(child)
            return
        for k in range(-child):
            target, blas_off, source = ints[index * 4 + 8 + k * 4: index * 4 + 11 + k * 4]
            seen.append(int(source))
            assert blas_off == (3 if source % 2 else 0)
            assert rec[target * 40 + 2] == source
            assert recf[target * 40 + 20 + 12] == source
    visit(0)
    assert sorted(seen) == list(range(n))
    assert t["instance_words"].size == n * 40


def test_tlas_refit_equals_rebuild_bounds(orc):
    rng = np.random.default_rng(2)
    n = 40
    tr = np.tile(np.eye(4).reshape(16), (n, 1))
    tr[:, 12:15] = rng.random((n, 3)) * 20
    lo, hi = np.zeros(3, np.float32), np.ones(3, np.float32)
    bounds = np.array([orc.instance_world_bounds(tr[i], lo, hi) for i in range(n)])
    inst = dict(blas_index=np.zeros(n, int), bounds=bounds, transform=tr, material_word=np.zeros(n, int))
    h = orc.build_tlas(inst, [0], [0])
    before = h.get()
    dirty = [3, 17]
    tr2 = tr.copy()
    tr2[dirty, 13] += 5.0
    bounds2 = np.array([orc.instance_world_bounds(tr2[i], lo, hi) for i in range(n)])
    inst2 = dict(inst, transform=tr2, bounds=bounds2)
    h.refit(inst2, dirty)
    after = h.get()
    f = after["words"].view(np.float32)
    ints = after["words"].view(np.int32)
    # every node bounds the instance boxes below it exactly (tight refit)

    def visit(index):
        child = ints[index * 4 + 3]
        if child >= 0:
            a, b = visit(index + 2), visit(child)
            exp = np.concatenate([np.minimum(a[:3], b[:3]), np.maximum(a[3:], b[3:])])
        else:
            srcs = [ints[index * 4 + 8 + k * 4 + 2] for k in range(-child)]
            exp = np.concatenate([bounds2[srcs, :3].min(0), bounds2[srcs, 3:].max(0)])
        act = np.array([f[index * 4 + k] for k in (0, 1, 2, 4, 5, 6)], np.float32)
        assert np.array_equal(act, exp.astype(np.float32))
        return act
    visit(0)
    assert not np.array_equal(before["words"], after["words"])


def test_world_bounds_translation_scale(orc):
    m = np.eye(4)
    m[0, 0] = 2.0
    m = m.T.reshape(16).copy()
    m[12:15] = [1, 2, 3]
    out = orc.instance_world_bounds(m, np.array([0, 0, 0], np.float32), np.array([1, 1, 1], np.float32))
    assert np.array_equal(out, np.array([1, 2, 3, 3, 3, 4], np.float32))


def _subtree_bounds(words, positions):
    """Brute force: each node's box = the bounds of every finite vertex of
    every triangle below it (independent of the refit's bottom-up order)."""
    w = np.asarray(words, np.uint32).reshape(-1, 4)
    code = w[:, 3].view(np.int32)
    out = {}

    def verts(row):
        c = code[row]
        if c < 0:
            return set(int(v) for v in w[row + 2:row + 2 - c, 0:3].reshape(-1))
        return verts(row + 2) | verts(int(c))

    stack = [0]
    while stack:
        row = stack.pop()
        vs = [v for v in verts(row) if np.all(np.isfinite(positions[v]))]
        p = positions[vs].astype(np.float64) if vs else np.zeros((0, 3))
        out[row] = (p.min(0) if vs else np.full(3, np.inf), p.max(0) if vs else np.full(3, -np.inf))
        if code[row] >= 0:
            stack += [row + 2, int(code[row])]
    return out


def test_refit_identity_and_subtree_bounds(orc):
    """refitBLAS (bvhAccel.ts:800-870): with unchanged positions the refit
    reproduces the build's words and surface area; after a deformation every
    node box equals the brute-force bounds of the finite vertices below it
    (an interior box taken from one child only, or a leaf box that keeps a
    NaN vertex, fails this)."""
    from mutracer_amd import workloads as W
    pos, idx = W.random_triangles(400, seed=7)
    ref = orc.build_blas(pos, idx, 8)
    w, area = orc.refit_blas(ref["words"], pos)
    assert np.array_equal(w, ref["words"])
    assert area == pytest.approx(ref["build_surface_area"], rel=1e-12)
    rng = np.random.default_rng(2)
    moved = (pos * np.array([1.5, 0.7, 1.0], np.float32) + rng.normal(0, 0.02, pos.shape)).astype(np.float32)
    moved[5] = np.nan
    w2, area2 = orc.refit_blas(ref["words"], moved)
    ww = w2.reshape(-1, 4)
    f = ww.view(np.float32)
    bounds = _subtree_bounds(w2, moved)
    for row, (lo, hi) in bounds.items():
        assert np.array_equal(f[row, 0:3], lo.astype(np.float32)), row
        assert np.array_equal(f[row + 1, 0:3], hi.astype(np.float32)), row
    # topology and leaf entries are untouched
    code = ww[:, 3].view(np.int32)
    orig = np.asarray(ref["words"], np.uint32).reshape(-1, 4)
    assert np.array_equal(ww[:, 3], orig[:, 3])
    leaf_rows = np.zeros(len(ww), bool)
    for row in bounds:
        if code[row] < 0:
            leaf_rows[row + 2:row + 2 - code[row]] = True
    assert np.array_equal(ww[leaf_rows], orig[leaf_rows])
    assert area2 > 0
