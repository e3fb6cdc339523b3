"""Oracle camera: pinhole rays and the thin lens (rayTrace.compute.ts:1188-1205,
sampleCircle :1148-1154), pinned by geometry rather than by its own formula:

- every lens origin lies on the lens disk (camera plane through the eye,
  radius = aperture);
- every lens ray of a pixel sample passes through the point where that
  sample's pinhole ray meets the focus plane (at focus_distance along the
  view direction);
- the lens point is the Shirley-Chiu concentric map of the CAMERA_LENS
  Sobol sample: radius = aperture * max(|p.x|, |p.y|) for p = 2u - 1, polar
  angle in the sector of the square's triangle that holds p;
- over many samples the lens points cover the disk uniformly.
"""
import math

import numpy as np
import pytest

import oracle
from mutracer_amd import scene

W, H = 64, 48


def _cam(aperture, focus):
    c = scene.Camera(eye=(1.0, 2.0, 9.0), target=(0.3, 0.5, 0.0), fov=40.0, aperture=aperture,
                     focus_distance=focus)
    return c, c.pack(W, H)


def _pixels(n, seed=0):
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(0, W, n), rng.integers(0, H, n), rng.integers(0, 4096, n)], 1).astype(np.uint32)


def _frame(c):
    _, vi = c.matrices(W, H)
    m = vi.reshape(4, 4).astype(np.float64)  # rows = columns of the view inverse
    return m[3, :3], m[0, :3], m[1, :3], -m[2, :3] / np.linalg.norm(m[2, :3])


def test_pinhole_unchanged_without_aperture():
    _, cam = _cam(0.0, 5.0)
    pix = _pixels(64)
    r = oracle.camera_rays(cam, W, H, pix)
    assert np.allclose(np.linalg.norm(r[:, 3:6], axis=1), 1.0, atol=1e-6)
    c, _ = _cam(0.0, 5.0)
    eye, _, _, fwd = _frame(c)
    # origins on the near plane, in front of the eye
    assert np.all((r[:, 0:3] - eye) @ fwd > 0.0)


@pytest.mark.parametrize("aperture,focus", [(0.2, 6.0), (0.05, 2.5)])
def test_thin_lens_geometry(aperture, focus):
    c, cam = _cam(aperture, focus)
    _, pin = _cam(0.0, focus)
    pix = _pixels(4096, seed=1)
    lens = oracle.camera_rays(cam, W, H, pix).astype(np.float64)
    hole = oracle.camera_rays(pin, W, H, pix).astype(np.float64)
    eye, x, y, fwd = _frame(c)
    off = lens[:, 0:3] - eye
    assert np.abs(off @ fwd).max() < 1e-5 * (1 + aperture)
    rad = np.linalg.norm(off, axis=1)
    assert rad.max() <= aperture * (1 + 1e-5)
    # focus point of the pinhole ray through the eye
    d = hole[:, 3:6]
    F = eye + d * (focus / (d @ fwd))[:, None]
    o, u = lens[:, 0:3], lens[:, 3:6]
    t = np.einsum("ij,ij->i", F - o, u)
    miss = np.linalg.norm(o + u * t[:, None] - F, axis=1)
    assert miss.max() < 2e-5 * focus


def test_concentric_map_of_lens_sample():
    aperture = 0.3
    c, cam = _cam(aperture, 4.0)
    pix = _pixels(2048, seed=2)
    lens = oracle.camera_rays(cam, W, H, pix).astype(np.float64)
    eye, x, y, _ = _frame(c)
    lx, ly = (lens[:, 0:3] - eye) @ x / aperture, (lens[:, 0:3] - eye) @ y / aperture
    u = np.array([oracle.structured_sample(0, px, py, W, idx, 1, 0, 0) for px, py, idx in pix], np.float64)
    p = 2.0 * u - 1.0
    r = np.maximum(np.abs(p[:, 0]), np.abs(p[:, 1]))
    assert np.allclose(np.hypot(lx, ly), r, atol=2e-5)
    theta = np.arctan2(ly, lx)
    big = r > 0.05
    right = (np.abs(p[:, 0]) > np.abs(p[:, 1])) & (p[:, 0] > 0) & big
    left = (np.abs(p[:, 0]) > np.abs(p[:, 1])) & (p[:, 0] < 0) & big
    top = (np.abs(p[:, 0]) <= np.abs(p[:, 1])) & (p[:, 1] > 0) & big
    bottom = (np.abs(p[:, 0]) <= np.abs(p[:, 1])) & (p[:, 1] < 0) & big
    q = math.pi / 4 + 1e-4
    assert np.all(np.abs(theta[right]) <= q)
    assert np.all(np.abs(np.abs(theta[left]) - math.pi) <= q)
    assert np.all(np.abs(theta[top] - math.pi / 2) <= q)
    assert np.all(np.abs(theta[bottom] + math.pi / 2) <= q)
    # the angle is linear in p within each triangle (pi/4 * p.y / p.x on the right)
    assert np.allclose(theta[right], math.pi / 4 * p[right, 1] / p[right, 0], atol=1e-5)


def test_lens_points_cover_disk_uniformly():
    aperture = 1.0
    c, cam = _cam(aperture, 4.0)
    n = 4096
    pix = np.stack([np.full(n, 17), np.full(n, 9), np.arange(n)], 1).astype(np.uint32)
    lens = oracle.camera_rays(cam, W, H, pix).astype(np.float64)
    eye, x, y, _ = _frame(c)
    off = np.stack([(lens[:, 0:3] - eye) @ x, (lens[:, 0:3] - eye) @ y], 1)
    r2 = (off ** 2).sum(1)
    assert abs(off[:, 0].mean()) < 0.01 and abs(off[:, 1].mean()) < 0.01
    assert abs(r2.mean() - 0.5) < 0.01          # E[r^2] = R^2 / 2 on the unit disk
    assert abs((r2 < 0.25).mean() - 0.25) < 0.01  # inner half-radius disk holds a quarter
