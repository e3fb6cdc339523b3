"""GPU binned-SAH build and TLAS build/refit vs the oracle: bit-exact trees.

The GPU builder must produce the reference builder's tree for the same
input: exported BVH2 words (bounds, pointers, leaf order, indices) and the
source triangle of every leaf record are compared word for word with the
oracle's sequential restatement of binnedSahBvh.ts / bvhAccel.ts.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mtx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from mutracer_amd import accel
    return accel


def _blas_case(name):
    from mutracer_amd import workloads as W
    if name.startswith("rand"):
        n = int(name[4:])
        return W.random_triangles(n, seed=n)
    if name == "sphere":
        return W.sphere_mesh(64, 32)
    if name == "coincident":
        p = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (7, 1))
        return p, np.arange(21, dtype=np.uint32)
    if name == "nan":
        p = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, np.nan, 0, 0, 3, 0, 0, 2, 1, 0], np.float32).reshape(-1, 3)
        return p, np.arange(6, dtype=np.uint32)
    if name == "single":
        return np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.arange(3, dtype=np.uint32)
    if name == "grid":  # many coplanar, axis-aligned (zero-extent axis) triangles
        xs, ys = np.meshgrid(np.arange(60, dtype=np.float32), np.arange(60, dtype=np.float32))
        v = np.stack([xs.ravel(), ys.ravel(), np.zeros(xs.size, np.float32)], 1)
        idx = []
        for j in range(59):
            for i in range(59):
                a = j * 60 + i
                idx += [a, a + 1, a + 60, a + 1, a + 61, a + 60]
        return v, np.array(idx, np.uint32)
    raise KeyError(name)


@pytest.mark.parametrize("case,bins", [("single", 8), ("nan", 8), ("coincident", 8), ("rand700", 8),
                                       ("rand5000", 8), ("rand5000", 32), ("sphere", 8), ("grid", 8),
                                       ("rand60000", 8), ("rand60000", 32), ("rand60000", 64)])
def test_blas_bit_exact(orc, mtx, case, bins):
    pos, idx = _blas_case(case)
    print(f"case={case} bins={bins} tris={idx.size // 3}", flush=True)
    ref = orc.build_blas(pos, idx, bins)
    b = mtx.Blas(pos, idx, bins)
    words, src = b.export()
    assert words.size == ref["words"].size
    assert np.array_equal(words, ref["words"]), f"first diff at {np.argmax(words != ref['words'])}"
    assert np.array_equal(src, ref["source_of_record"])
    info = b.info()
    assert info["max_depth"] == ref["max_depth"]
    assert info["max_leaf"] == ref["max_leaf"]


def test_blas_device_input_matches_host_input(orc, mtx):
    import torch
    pos, idx = _blas_case("rand5000")
    a = mtx.Blas(pos, idx, 8).export()[0]
    b = mtx.Blas(torch.from_numpy(pos).cuda(), torch.from_numpy(idx.astype(np.int32)).cuda(), 8).export()[0]
    assert np.array_equal(a, b)


def _instanced(n_inst, seed=1):
    from mutracer_amd import workloads as W
    rng = np.random.default_rng(seed)
    meshes = [W.sphere_mesh(24, 12), W.random_triangles(3000, seed=9)]
    mesh_idx = rng.integers(0, 2, n_inst)
    tr = []
    for i in range(n_inst):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        tr.append(W.trs(rng.random(3) * 20 - 10, q, float(rng.lognormal(0, 0.3))))
    return meshes, mesh_idx, np.array(tr)


@pytest.mark.parametrize("bins,per_inst", [(64, True), (8, False)])
def test_tlas_matches_oracle(orc, mtx, bins, per_inst):
    meshes, mesh_idx, tr = _instanced(40)
    mats = np.arange(40) % 3
    scene = orc.assemble_scene([dict(positions=m[0], indices=m[1]) for m in meshes],
                               dict(mesh=mesh_idx, transform=tr, material=mats), tlas_bins=bins,
                               per_instance_bounds=per_inst)
    blases = [mtx.Blas(m[0], m[1]) for m in meshes]
    acc = mtx.Accel(blases, mesh_idx, tr, material=mats, tlas_bins=bins, per_instance_bounds=per_inst)
    tw, iw = acc.export_tlas()
    ref_t = scene.keep[0].copy()
    ref_i = scene.keep[2].copy().reshape(-1, 40)
    # BLAS offsets are buffer addresses in the resident layout: compare the rest.
    ints = ref_t.view(np.int32)
    got = tw.view(np.int32).copy()
    stack = [0]
    while stack:
        node = stack.pop()
        c = ints[node * 4 + 3]
        if c >= 0:
            stack += [c, node + 2]
            continue
        for k in range(-c):
            got[node * 4 + 8 + k * 4 + 1] = ints[node * 4 + 8 + k * 4 + 1]
    assert np.array_equal(got, ints)
    gi = iw.reshape(-1, 40)
    assert np.array_equal(gi[:, 2:], ref_i[:, 2:])


def test_tlas_refit_matches_oracle(orc, mtx):
    meshes, mesh_idx, tr = _instanced(30, seed=4)
    mats = np.zeros(30, int)
    scene = orc.assemble_scene([dict(positions=m[0], indices=m[1]) for m in meshes],
                               dict(mesh=mesh_idx, transform=tr, material=mats))
    blases = [mtx.Blas(m[0], m[1]) for m in meshes]
    acc = mtx.Accel(blases, mesh_idx, tr, material=mats)
    tr2 = tr.copy()
    dirty = [2, 11, 29]
    tr2[dirty, 12:15] += np.array([3.0, -2.0, 1.5])
    acc.refit(tr2, dirty)
    # oracle: world bounds from the BLAS root boxes, then full refit
    roots = []
    for m in meshes:
        w = orc.build_blas(*m)["words"]
        roots.append((w[0:3].view(np.float32).copy(), w[4:7].view(np.float32).copy()))
    bounds = np.array([orc.instance_world_bounds(tr2[i], *roots[mesh_idx[i]]) for i in range(30)])
    scene.tlas_handle.refit(dict(blas_index=mesh_idx, bounds=bounds, transform=tr2,
                                 material_word=np.zeros(30, np.uint32)), dirty)
    ref = scene.tlas_handle.get()
    tw, iw = acc.export_tlas()
    ints = ref["words"].view(np.int32)
    got = tw.view(np.int32).copy()
    f_ref = ref["words"].view(np.float32)
    f_got = tw.view(np.float32)
    stack = [0]
    while stack:
        node = stack.pop()
        assert np.array_equal(f_got[node * 4:node * 4 + 3], f_ref[node * 4:node * 4 + 3])
        assert np.array_equal(f_got[node * 4 + 4:node * 4 + 7], f_ref[node * 4 + 4:node * 4 + 7])
        c = ints[node * 4 + 3]
        if c >= 0:
            stack += [c, node + 2]
            continue
        for k in range(-c):
            got[node * 4 + 8 + k * 4 + 1] = ints[node * 4 + 8 + k * 4 + 1]
    assert np.array_equal(got, ints)
    gi = iw.reshape(-1, 40)
    ri = ref["instance_words"].reshape(-1, 40)
    assert np.array_equal(gi[:, 4:36], ri[:, 4:36])
