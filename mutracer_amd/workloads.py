"""Synthetic workloads of the benchmark configurations (host-side numpy).

All generators are deterministic functions of their seed (numpy PCG64).
Sizes are parameters so that tests can run scaled-down versions of the same
construction.
"""
from __future__ import annotations

import math

import numpy as np


# ---------------------------------------------------------------- config 3
def random_triangles(n, seed=3, mu=-5.0, sigma=0.5):
    """n independent triangles: centroid U[0,1]^3, three vertices at
    centroid + lognormal(mu, sigma) edge length along random unit directions."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.random((n, 3))
    dirs = rng.normal(size=(n, 3, 3))
    dirs /= np.linalg.norm(dirs, axis=2, keepdims=True)
    length = rng.lognormal(mu, sigma, size=(n, 3, 1))
    v = c[:, None, :] + dirs * length
    positions = v.reshape(-1, 3).astype(np.float32)
    indices = np.arange(3 * n, dtype=np.uint32)
    return positions, indices


def pinhole_rays(res, eye=(0.5, 0.5, -1.2), target=(0.5, 0.5, 0.5), fov_deg=45.0):
    """res x res pinhole primaries (float32 rows: o.xyz, tmax, d.xyz, 0)."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    t = math.tan(math.radians(fov_deg) * 0.5)
    j, i = np.meshgrid(np.arange(res), np.arange(res), indexing="ij")
    u = ((i + 0.5) / res * 2 - 1) * t
    v = (1 - (j + 0.5) / res * 2) * t
    d = fwd[None, None, :] + u[..., None] * right + v[..., None] * up
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    rays = np.zeros((res * res, 8), np.float32)
    rays[:, 0:3] = eye
    rays[:, 3] = 9999.0
    rays[:, 4:7] = d.reshape(-1, 3)
    return rays


def hemisphere_bounces(positions_world, normals, count, seed=0, eps=1e-4):
    """Cosine-hemisphere bounce rays about the geometric normals of hit points,
    round-robin over the hits, offset along the normal."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = positions_world.shape[0]
    idx = np.arange(count) % m
    p = positions_world[idx].astype(np.float64)
    n = normals[idx].astype(np.float64)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    u1, u2 = rng.random(count), rng.random(count)
    r = np.sqrt(u1)
    phi = 2 * math.pi * u2
    local = np.stack([r * np.cos(phi), r * np.sin(phi), np.sqrt(np.maximum(0.0, 1 - u1))], 1)
    a = np.where(np.abs(n[:, 0:1]) > 0.9, [[0.0, 1.0, 0.0]], [[1.0, 0.0, 0.0]])
    t = np.cross(n, a)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(n, t)
    d = local[:, 0:1] * t + local[:, 1:2] * b + local[:, 2:3] * n
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((count, 8), np.float32)
    rays[:, 0:3] = p + n * eps
    rays[:, 3] = 9999.0
    rays[:, 4:7] = d
    return rays


def identity_transform():
    return np.eye(4).reshape(16)


def trs(translation, quat_xyzw, scale):
    """Column-major 4x4 from translation, rotation quaternion and uniform scale."""
    x, y, z, w = quat_xyzw
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    M = np.eye(4)
    M[:3, :3] = R * scale
    M[:3, 3] = translation
    return M.T.reshape(16).copy()  # column-major


def sphere_mesh(seg=32, rings=16, radius=1.0):
    verts, idx = [], []
    for r in range(rings + 1):
        th = math.pi * r / rings
        for s in range(seg + 1):
            ph = 2 * math.pi * s / seg
            verts.append([radius * math.sin(th) * math.cos(ph), radius * math.cos(th),
                          radius * math.sin(th) * math.sin(ph)])
    for r in range(rings):
        for s in range(seg):
            a, b = r * (seg + 1) + s, (r + 1) * (seg + 1) + s
            idx += [a, b, a + 1, b, b + 1, a + 1]
    return np.array(verts, np.float32), np.array(idx, np.uint32)
