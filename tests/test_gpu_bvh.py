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


@pytest.mark.parametrize("case", ["rand5000", "sphere", "nan"])
def test_blas_refit_matches_oracle(orc, mtx, case):
    """refitBLAS on the GPU (mtx_blas_refit) against the oracle's restatement
    on the GPU tree's own words: identical words after an identity refit and
    after a deformation (scale, shear, noise, one vertex made non-finite);
    the rebuild advice follows the surface-area ratio. An identity refit
    reproduces the build's boxes up to the sign of zero bounds, which the
    reference's build and refit derive along different paths (the oracle's
    own identity refit of the sphere flips five zero words)."""
    pos, idx = _blas_case(case)
    b = mtx.Blas(pos, idx, 8)
    words0, _ = b.export()
    assert b.refit(pos) is False
    w1, _ = b.export()
    ref1, _ = orc.refit_blas(words0, pos)
    assert np.array_equal(w1, ref1)
    moved_bits = w1 != words0
    assert np.all((w1[moved_bits] | words0[moved_bits]) == 0x80000000)
    rng = np.random.default_rng(4)
    moved = pos.astype(np.float64) @ np.array([[1.3, 0.2, 0.0], [0.0, 0.8, 0.1], [0.0, 0.0, 1.1]])
    moved = (moved + rng.normal(0, 0.01, pos.shape)).astype(np.float32)
    if case != "nan":
        moved[3] = np.inf
    advised = b.refit(moved)
    got, _ = b.export()
    ref, area = orc.refit_blas(words0, moved)
    assert np.array_equal(got, ref)
    _, build_area = orc.refit_blas(words0, pos)
    assert advised == (area > 2.0 * build_area)
    big = (pos * 3.0).astype(np.float32)
    assert b.refit(big) is True  # nine times the build's surface area


def test_refit_traversal_bit_exact(orc, mtx):
    """Closest hits through a refitted BLAS equal the oracle's traversal of the
    refitted words, bit for bit."""
    import torch
    from mutracer_amd import workloads as W
    pos, idx = W.random_triangles(20000, seed=9)
    blas = mtx.Blas(pos, idx, 8)
    rng = np.random.default_rng(1)
    moved = (pos + rng.normal(0, 0.003, pos.shape)).astype(np.float32)
    blas.refit(moved)
    acc = mtx.Accel([blas], [0], [W.identity_transform()], material=[0], double_sided=[1])
    words, src = blas.export()
    tw, iw = acc.export_tlas()
    scene = orc.TraceScene(tw, words, iw, moved, np.ones(1, np.uint8))
    rays = W.pinhole_rays(96)
    r = torch.from_numpy(rays).cuda()
    h = torch.zeros((rays.shape[0], 8), dtype=torch.int32, device="cuda")
    acc.trace_closest(r, h, double_sided=True)
    torch.cuda.synchronize()
    got = h.cpu().numpy().view(np.uint32)
    ref = scene.closest(rays, double_sided=True)
    hit = ref[:, 0].astype(bool)
    assert np.array_equal(got[:, 0].astype(bool), hit)
    assert np.array_equal(got[hit, 2], src[ref[hit, 2]])
    assert np.array_equal(got[hit, 3], ref[hit, 3])
    assert hit.mean() > 0.1


def test_scene_update_mesh_render(orc):
    """mtx_scene_update_mesh (BLAS refit + TLAS refit) on a committed scene:
    the render of the deformed scene equals the oracle's render of the same
    deformed geometry (the oracle builds its own tree; its closest hits agree
    except at exact-t ties) within the accuracy bar relMSE <= 1e-4, and an
    identity update leaves the image bit-identical."""
    from mutracer_amd import render, scene
    sc = scene.instanced_field(n_instances=6, freq=8, env_size=(64, 32), seed=2)
    rs = render.RenderScene(sc)
    W_, H_, spp = 40, 24, 4
    kw = dict(bounces=6, direct_light_candidates=2, max_samples=spp)
    t = render.RenderTarget(rs, W_, H_, **kw)
    t.render(spp)
    before, _ = t.read_accumulators()
    m = 2
    pos, nrm = sc.meshes[m]["positions"].copy(), sc.meshes[m]["normals"].copy()
    assert rs.update_mesh(m, pos, nrm) is False
    t.restart_samples()
    t.render(spp)
    again, _ = t.read_accumulators()
    assert np.array_equal(again, before)
    moved = (pos * np.float32(1.15)).astype(np.float32)
    rs.update_mesh(m, moved, nrm)
    t.restart_samples()
    t.render(spp)
    got, _ = t.read_accumulators()
    ref = orc.OracleRenderer(rs.scene, W_, H_, bounces=6, candidates=2, max_samples=spp).render(0, spp)
    c_got, c_ref = got[..., :3], ref[..., :3]
    rel = float(np.mean((c_got - c_ref) ** 2 / (c_ref ** 2 + 1e-2)))
    assert rel <= 1e-4, rel
    assert not np.array_equal(got, before)
