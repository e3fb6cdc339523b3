"""GPU two-level traversal vs the oracle: bit-exact closest-hit IDs and t.

For identical f32 rays the HIP traversal must return the same hit flag,
source instance, source triangle, t bits and barycentric bits as the
oracle's reference-order traversal (intersect.wgsl.ts getBVHTraverseCode).
Any-hit occlusion and the per-ray visit counters are compared as well.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from mutracer_amd import accel, workloads
    return torch, accel, workloads


def _gpu_closest(torch, acc, rays, double_sided=False):
    r = torch.from_numpy(rays).cuda()
    h = torch.zeros((rays.shape[0], 8), dtype=torch.int32, device="cuda")
    acc.trace_closest(r, h, double_sided=double_sided)
    torch.cuda.synchronize()
    return h.cpu().numpy().view(np.uint32)


def _compare(scene, ref_hits, gpu_hits, orc):
    hit, inst, tri, tbits = orc.hits_to_ids(scene, ref_hits)
    g = gpu_hits
    assert np.array_equal(g[:, 0].astype(bool), hit), "hit flags differ"
    assert np.array_equal(g[hit, 1].view(np.int32), inst[hit]), "instance ids differ"
    assert np.array_equal(g[hit, 2].view(np.int32), tri[hit]), "triangle ids differ"
    assert np.array_equal(g[hit, 3], tbits[hit]), "t bits differ"
    assert np.array_equal(g[hit, 4:7], ref_hits[hit, 4:7]), "barycentric bits differ"
    return hit.mean()


def _c3_scene(orc, accel, W, ntris, res, bins=32):
    pos, idx = W.random_triangles(ntris, seed=3)
    scene = orc.assemble_scene([dict(positions=pos, indices=idx)],
                               dict(mesh=[0], transform=[W.identity_transform()], material=[0]),
                               blas_bins=bins, material_double_sided=[1])
    blas = accel.Blas(pos, idx, bins)
    acc = accel.Accel([blas], [0], [W.identity_transform()], material=[0], double_sided=[1])
    return pos, idx, scene, acc, blas


@pytest.mark.parametrize("ntris,res", [(2000, 64), (100000, 192)])
def test_microbench_scaled_bit_exact(orc, env, ntris, res):
    torch, accel, W = env
    print(f"microbench ntris={ntris} res={res}", flush=True)
    pos, idx, scene, acc, _ = _c3_scene(orc, accel, W, ntris, res)
    prim = W.pinhole_rays(res)
    ref = scene.closest(prim)
    got = _gpu_closest(torch, acc, prim)
    frac = _compare(scene, ref, got, orc)
    assert frac > (0.2 if ntris >= 100000 else 0.005)
    # bounce rays from the primary hits, about the geometric normal
    hit = ref[:, 0].astype(bool)
    t = ref[hit, 3].view(np.float32)
    p = prim[hit, 0:3] + prim[hit, 4:7] * t[:, None]
    tri = scene.record_triangle[ref[hit, 2]]
    v = pos[idx.reshape(-1, 3)[tri]]
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    n *= np.where((n * prim[hit, 4:7]).sum(1, keepdims=True) > 0, -1, 1)
    bounce = W.hemisphere_bounces(p, n, prim.shape[0], seed=0)
    ref_b = scene.closest(bounce)
    got_b = _gpu_closest(torch, acc, bounce)
    _compare(scene, ref_b, got_b, orc)


def test_single_sided_culling_and_any_hit(orc, env):
    torch, accel, W = env
    pos, idx = W.sphere_mesh(40, 20, radius=0.4)
    pos = pos + np.float32(0.5)
    scene = orc.assemble_scene([dict(positions=pos, indices=idx)],
                               dict(mesh=[0], transform=[W.identity_transform()], material=[0]),
                               material_double_sided=[0])
    acc = accel.Accel([accel.Blas(pos, idx)], [0], [W.identity_transform()], material=[0], double_sided=[0])
    rays = W.pinhole_rays(96)
    # from inside the sphere too: back faces are culled for single-sided materials
    inner = rays.copy()
    inner[:, 0:3] = 0.5
    for rr in (rays, inner):
        _compare(scene, scene.closest(rr), _gpu_closest(torch, acc, rr), orc)
        _compare(scene, scene.closest(rr, double_sided=True), _gpu_closest(torch, acc, rr, True), orc)
    short = rays.copy()
    short[:, 3] = np.linspace(0.0, 2.5, short.shape[0], dtype=np.float32)
    r = torch.from_numpy(short).cuda()
    occ = torch.zeros(short.shape[0], dtype=torch.int32, device="cuda")
    acc.trace_any(r, occ)
    torch.cuda.synchronize()
    assert np.array_equal(occ.cpu().numpy().astype(np.uint32), scene.any(short))


def test_instanced_scene_bit_exact(orc, env):
    torch, accel, W = env
    rng = np.random.default_rng(7)
    meshes = [W.sphere_mesh(24, 12), W.random_triangles(4000, seed=5), W.sphere_mesh(8, 4)]
    n = 48
    mesh_idx = rng.integers(0, 3, n)
    tr = []
    for i in range(n):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        tr.append(W.trs(rng.random(3) * 6 - 3, q, float(rng.lognormal(0, 0.3))))
    tr = np.array(tr)
    mats = rng.integers(0, 2, n)
    ds = [0, 1]
    scene = orc.assemble_scene([dict(positions=m[0], indices=m[1]) for m in meshes],
                               dict(mesh=mesh_idx, transform=tr, material=mats), material_double_sided=ds)
    blases = [accel.Blas(m[0], m[1]) for m in meshes]
    acc = accel.Accel(blases, mesh_idx, tr, material=mats, double_sided=[ds[m] for m in mats])
    rays = W.pinhole_rays(160, eye=(0, 0, -9), target=(0, 0, 0), fov_deg=50)
    frac = _compare(scene, scene.closest(rays), _gpu_closest(torch, acc, rays), orc)
    assert frac > 0.05
    # counters equal the oracle's instrumented traversal
    r = torch.from_numpy(rays).cuda()
    cnt = torch.zeros((rays.shape[0], 3), dtype=torch.int32, device="cuda")
    acc.trace_counters(r, cnt)
    torch.cuda.synchronize()
    _, st = scene.closest(rays, stats=True)
    assert np.array_equal(cnt.cpu().numpy().astype(np.uint32), st)
