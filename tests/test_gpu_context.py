"""Persistent contexts over raw scene descriptors, on the GPU.

A committed descriptor renders the scene it describes: the context's images
equal the oracle's FP64 sums divided by the completed sample count (the
reference's read_pixels, native/host/src/lib.rs:343-364), bit for bit, and
equal a render target of the same Scene. Jobs follow the one-job-per-context
contract of native/include/mutracer-context.h:146-187 (submit / poll / info /
cancel, draining cancellation, resume), and a native C host written against
include/mtx.h alone runs the same flow in a child process.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def mods():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from mutracer_amd import abi, context, descriptor, render, scene
    return abi, context, descriptor, render, scene


def _settings(context, spp, bounces=8, candidates=1, pattern=0, seed=0):
    return context.settings_default(seed=seed, bounces=bounces, direct_light_candidates=candidates,
                                    max_samples=spp, sampling_pattern=pattern)


def _expected(sums, spp, base):
    return (sums[..., base:base + 4] / spp).astype(np.float32)


def test_cornell_descriptor_matches_oracle(mods, orc):
    abi, context, descriptor, render, scene = mods
    sc = scene.cornell_box()
    W, H, spp = 40, 32, 5
    data = descriptor.encode(sc, W, H)
    ctx = context.RenderContext()
    ctx.commit(data, _settings(context, spp))
    ctx.render(spp)
    info = ctx.info()
    assert info == dict(width=W, height=H, samples=spp, state="committed", last_commit="full")
    ref = orc.OracleRenderer(descriptor.f32_scene(sc), W, H, bounces=8, candidates=1, max_samples=spp).render(0, spp)
    for kind, base in (("hdr", 0), ("albedo", 4), ("normal", 8)):
        assert np.array_equal(ctx.read_pixels(kind), _expected(ref, spp, base)), kind


def test_instanced_field_descriptor_matches_target(mods):
    abi, context, descriptor, render, scene = mods
    sc = scene.instanced_field(n_instances=9, freq=10, env_size=(128, 64))
    W, H, spp = 48, 27, 4
    ctx = context.RenderContext()
    ctx.commit(descriptor.encode(sc, W, H), _settings(context, spp, bounces=12, candidates=4))
    ctx.render(spp)
    t = render.RenderTarget(render.RenderScene(descriptor.f32_scene(sc)), W, H, bounces=12,
                            direct_light_candidates=4, max_samples=spp)
    t.render(spp)
    sums, counts = t.read_accumulators()
    assert np.all(counts == spp)
    assert np.array_equal(ctx.read_pixels("hdr"), _expected(sums, spp, 0))


def test_job_poll_cancel_resume(mods):
    abi, context, descriptor, render, scene = mods
    sc = scene.cornell_box()
    data = descriptor.encode(sc, 32, 24)
    a, b = context.RenderContext(), context.RenderContext()
    st = _settings(context, 16)
    a.commit(data, st)
    b.commit(data, st)
    job = a.submit(16)
    assert job.info() == dict(state="queued", requested=16, submitted=0, completed=0, base=0)
    assert a.info()["state"] == "job_active"
    with pytest.raises(abi.MtxError) as e:
        a.render(1)
    assert e.value.code == abi.MTX_INVALID_STATE
    job.poll(4)
    assert job.info()["state"] == "running" and job.info()["submitted"] == 4
    job.cancel()
    assert job.info()["state"] in ("cancelling", "cancelled")
    while job.info()["state"] == "cancelling":
        job.poll(0)
    assert job.info() == dict(state="cancelled", requested=16, submitted=4, completed=4, base=0)
    assert a.info()["samples"] == 4
    job = a.submit(12)
    while job.info()["state"] != "complete":
        job.poll(5)
    assert job.info()["base"] == 4 and job.info()["completed"] == 12
    b.render(16)
    for kind in ("hdr", "normal", "albedo"):
        assert np.array_equal(a.read_pixels(kind), b.read_pixels(kind)), kind


def test_failure_marks_context_failed(mods):
    abi, context, descriptor, render, scene = mods
    data = descriptor.encode(scene.cornell_box(), 16, 16)
    ctx = context.RenderContext()
    ctx.commit(data)
    ctx.render(1)
    ctx.inject_failure()
    with pytest.raises(abi.MtxError) as e:
        ctx.render(1)
    assert e.value.code == abi.MTX_CONTEXT_FAILED
    assert ctx.info()["state"] == "failed"
    for call in (lambda: ctx.render(1), lambda: ctx.read_pixels(), lambda: ctx.reset(), lambda: ctx.submit(1)):
        with pytest.raises(abi.MtxError) as e:
            call()
        assert e.value.code == abi.MTX_INVALID_STATE
    ctx.commit(data)
    assert ctx.info()["state"] == "committed"
    ctx.render(2)
    assert ctx.info()["samples"] == 2
    # A failed poll marks the job failed too.
    job = ctx.submit(4)
    ctx.inject_failure()
    with pytest.raises(abi.MtxError) as e:
        job.poll(2)
    assert e.value.code == abi.MTX_CONTEXT_FAILED
    assert job.info()["state"] == "failed"
    ctx.commit(data)
    ctx.destroy()


def test_incremental_camera_and_transform_commits(mods):
    abi, context, descriptor, render, scene = mods
    sc = scene.cornell_box()
    W, H, spp = 32, 24, 3
    ctx = context.RenderContext()
    st = _settings(context, spp)
    ctx.commit(descriptor.encode(sc, W, H), st)
    ctx.render(spp)
    # Camera-only edit: incremental, samples restart, image = fresh commit's.
    moved = descriptor.f32_scene(sc)
    moved.camera = scene.Camera(eye=(0.2, 1.1, 3.8), target=(0.0, 0.9, 0.0), fov=41.0)
    data = descriptor.encode(moved, W, H)
    ctx.commit(data, st)
    assert ctx.info()["last_commit"] == "incremental" and ctx.info()["samples"] == 0
    ctx.render(spp)
    fresh = context.RenderContext()
    fresh.commit(data, st)
    fresh.render(spp)
    assert np.array_equal(ctx.read_pixels(), fresh.read_pixels())
    # Transform-only edit (one box moved): the TLAS is refitted; the image
    # equals a render target whose scene had the same refit.
    edited = descriptor.f32_scene(moved)
    t = edited.transforms[5].copy()
    t[12] += 0.05
    edited.transforms[5] = t
    ctx.commit(descriptor.encode(edited, W, H), st)
    assert ctx.info()["last_commit"] == "incremental"
    ctx.render(spp)
    rs = render.RenderScene(descriptor.f32_scene(moved))
    rs.set_transforms(np.array(descriptor.f32_scene(edited).transforms))
    tgt = render.RenderTarget(rs, W, H, bounces=8, direct_light_candidates=1, max_samples=spp, camera=moved.camera)
    tgt.render(spp)
    sums, _ = tgt.read_accumulators()
    assert np.array_equal(ctx.read_pixels(), _expected(sums, spp, 0))


def test_c_host_job_flow(mods, orc, tmp_path):
    """include/mtx.h compiled by gcc into a native host, run as a fresh child
    process; its resumed-job image equals the oracle's."""
    abi, context, descriptor, render, scene = mods
    sc = scene.cornell_box()
    W, H, spp = 32, 32, 12
    desc = tmp_path / "cornell.mtscene"
    desc.write_bytes(descriptor.encode(sc, W, H))
    exe = tmp_path / "job_host"
    lib = os.path.join(ROOT, "mutracer_amd")
    subprocess.run(["gcc", "-std=c11", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c_host", "job_host.c"), "-L", lib, "-lmtx",
                    f"-Wl,-rpath,{lib}", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), str(desc), str(tmp_path), str(spp)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("PASS")
    ref = orc.OracleRenderer(descriptor.f32_scene(sc), W, H, bounces=8, candidates=1, max_samples=spp).render(0, spp)
    for kind, base in (("hdr", 0), ("albedo", 4), ("normal", 8)):
        got = np.fromfile(tmp_path / f"{kind}.rgba32f", np.float32).reshape(H, W, 4)
        assert np.array_equal(got, _expected(ref, spp, base)), kind


@pytest.mark.parametrize("which", ["textured_room", "mask_checker_caster", "mask_checker_panes"])
def test_textured_descriptor_matches_target(mods, which):
    """The textured room through the descriptor (texture records, UV sets,
    texture slots in the basic record and in extensions) renders exactly as
    a render target of the same Scene; so do the mask-material scenes (the
    descriptor's alpha mode, cutoff, opacity and alpha map resolve to the
    same packed coverage)."""
    abi, context, descriptor, render, scene = mods
    sc = {"textured_room": scene.textured_room,
          "mask_checker_caster": lambda: scene.alpha_caster("checker"),
          "mask_checker_panes": lambda: scene.masked_panes(2, checker=True)}[which]()
    W, H, spp = 40, 30, 4
    ctx = context.RenderContext()
    ctx.commit(descriptor.encode(sc, W, H), _settings(context, spp, bounces=8, candidates=2))
    ctx.render(spp)
    t = render.RenderTarget(render.RenderScene(descriptor.f32_scene(sc)), W, H, bounces=8,
                            direct_light_candidates=2, max_samples=spp)
    t.render(spp)
    sums, counts = t.read_accumulators()
    assert np.all(counts == spp)
    for kind, base in (("hdr", 0), ("albedo", 4), ("normal", 8)):
        assert np.array_equal(ctx.read_pixels(kind), _expected(sums, spp, base)), kind
