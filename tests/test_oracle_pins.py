"""Independent pins of two oracle restatements that the traversal tests do not fix.

- ComputeBarycentrics (Include/TraverseBVHStackless.glsl:156-172) and its call
  in the scene loop (:300-315): vertices and edge midpoints of a known triangle
  against a float64 solve, random interior points against the same solve, and
  the reconstruction u*A + v*B + w*C of every hit point of a real walk.  A v/w
  swap, a wrong vertex order or the wrong point (origin instead of hit) fails.
- GetRayDirectionAt + main() (Intersectors/TraverseBVHStackless.glsl:143-148,
  371-388): the pixel-corner uv (no half-pixel offset), y growing upward, the
  half-FOV corner directions and the origin from the inverse view's last
  column, for a symmetric perspective and a rotated, translated camera.
"""
import math

import numpy as np
import pytest

from oracle_cases import random_rays


def _solve64(p, a, b, c):
    """(u, v, w) with u*a + v*b + w*c = p, u + v + w = 1, by float64 least squares."""
    a, b, c, p = (np.asarray(x, np.float64) for x in (a, b, c, p))
    m = np.stack([b - a, c - a], axis=1)
    vw, *_ = np.linalg.lstsq(m, p - a, rcond=None)
    return np.array([1.0 - vw[0] - vw[1], vw[0], vw[1]])


TRI = (np.array([1.5, -0.25, 2.0], np.float32), np.array([4.0, 1.0, 2.5], np.float32),
       np.array([2.0, 3.5, -1.0], np.float32))


@pytest.mark.parametrize("point,expect", [
    ("a", (1, 0, 0)), ("b", (0, 1, 0)), ("c", (0, 0, 1)),
    ("ab", (0.5, 0.5, 0)), ("bc", (0, 0.5, 0.5)), ("ca", (0.5, 0, 0.5)),
])
def test_barycentrics_vertices_and_midpoints(orc, point, expect):
    a, b, c = TRI
    pts = {"a": a, "b": b, "c": c, "ab": (a + b) / 2, "bc": (b + c) / 2, "ca": (c + a) / 2}
    p = pts[point].astype(np.float32)
    got = orc.barycentrics(p, a, b, c)
    assert np.allclose(_solve64(p, a, b, c), expect, atol=1e-12)
    assert np.abs(got - np.array(expect)).max() < 1e-6, (point, got)


def test_barycentrics_random_vs_float64(orc):
    rng = np.random.default_rng(11)
    worst = 0.0
    for _ in range(2000):
        a, b, c = rng.normal(size=(3, 3)).astype(np.float32) * 4
        n = np.cross(b.astype(np.float64) - a, c.astype(np.float64) - a)
        e = max(np.linalg.norm(b - a), np.linalg.norm(c - a), np.linalg.norm(c - b))
        if np.linalg.norm(n) < 0.05 * e * e:          # slivers are ill-conditioned
            continue
        wts = rng.dirichlet([1, 1, 1])
        p = (wts[0] * a.astype(np.float64) + wts[1] * b + wts[2] * c).astype(np.float32)
        got = orc.barycentrics(p, a, b, c)
        ref = _solve64(p, a, b, c)
        worst = max(worst, np.abs(got - ref).max())
    assert worst < 1e-4


def test_barycentrics_reconstruct_hit_points(crt, orc):
    """The scene loop's (u, v, w) weight the hit triangle's vertices A, B, C in order."""
    import oracle
    from candela_amd import scenes
    for fmt in (0, 1):
        b = crt.build_bvh(fmt, scenes.boxes_scene())
        ent = oracle.identity_entity(b.nodes.size)
        rays = random_rays(3000, [-4, 0.5, -4], [20, 8, 20], seed=21)
        h = orc.trace(fmt, b.nodes, b.triangles, b.vertices, ent, 0, rays)
        hit = (h["t"] > 0) & (h["tri"] >= 0)
        assert hit.mean() > 0.3
        tri = b.triangles["packed"][h["tri"][hit]]
        pos = b.vertices["position"][:, :3].astype(np.float64)
        A, B, C = pos[tri[:, 0]], pos[tri[:, 1]], pos[tri[:, 2]]
        uvw = np.stack([h["u"][hit], h["v"][hit], h["w"][hit]], axis=1).astype(np.float64)
        rec = uvw[:, :1] * A + uvw[:, 1:2] * B + uvw[:, 2:3] * C
        p = rays["org"][hit].astype(np.float64) + rays["dir"][hit].astype(np.float64) * h["t"][hit, None]
        err = np.linalg.norm(rec - p, axis=1)
        assert err.max() < 1e-4, err.max()
        assert (uvw >= -1e-4).all() and (uvw <= 1 + 1e-4).all()
        # swapping v and w (or B and C) moves the point on every non-degenerate hit
        swapped = uvw[:, :1] * A + uvw[:, 2:3] * B + uvw[:, 1:2] * C
        moved = np.linalg.norm(swapped - p, axis=1) > 1e-3
        assert moved.mean() > 0.9


# --- primary rays --------------------------------------------------------------------

def _inv_perspective(fovy_deg, aspect, near=0.1, far=1000.0):
    """Inverse of the OpenGL perspective matrix (glm::perspective), row-major, float64."""
    f = 1.0 / math.tan(math.radians(fovy_deg) / 2)
    p = np.zeros((4, 4))
    p[0, 0] = f / aspect
    p[1, 1] = f
    p[2, 2] = (far + near) / (near - far)
    p[2, 3] = 2 * far * near / (near - far)
    p[3, 2] = -1.0
    return np.linalg.inv(p)


def _colmajor(m):
    return np.ascontiguousarray(np.asarray(m, np.float32).T.reshape(16))


def _norm(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def test_primary_rays_identity_view(orc):
    W, H, fovy = 64, 32, 60.0
    aspect = W / H
    ip = _inv_perspective(fovy, aspect)
    rays = orc.primary_rays(W, H, _colmajor(np.eye(4)), _colmajor(ip), tmax=123.0)
    d = rays["dir"].reshape(H, W, 3).astype(np.float64)
    t = math.tan(math.radians(fovy) / 2)
    # pixel (W/2, H/2) is the centre of the clip square: straight down -z
    assert np.abs(d[H // 2, W // 2] - [0, 0, -1]).max() < 1e-7
    # pixel (0, 0) is the bottom-left corner at half the FOV (corner uv = 0, no +0.5)
    assert np.abs(d[0, 0] - _norm([-t * aspect, -t, -1])).max() < 1e-6
    # last pixel: uv = (W-1)/W, (H-1)/H
    cx, cy = 2 * (W - 1) / W - 1, 2 * (H - 1) / H - 1
    assert np.abs(d[H - 1, W - 1] - _norm([cx * t * aspect, cy * t, -1])).max() < 1e-6
    # every pixel against the float64 closed form
    xs, ys = np.meshgrid(np.arange(W) / W * 2 - 1, np.arange(H) / H * 2 - 1)
    ref = _norm(np.stack([xs * t * aspect, ys * t, -np.ones_like(xs)], axis=-1))
    assert np.abs(d - ref).max() < 1e-6
    # y grows upward along rows, x to the right along columns
    assert (np.diff(d[:, W // 2, 1]) > 0).all() and (np.diff(d[H // 2, :, 0]) > 0).all()
    assert (rays["org"] == 0).all() and (rays["tmax"] == np.float32(123.0)).all()
    assert np.array_equal(rays["id"], np.arange(W * H, dtype=np.uint32))


def test_primary_rays_rotated_translated_view(orc):
    W, H, fovy = 48, 40, 75.0
    aspect = W / H
    ip = _inv_perspective(fovy, aspect)
    # camera at (3, -2, 7) looking down +x: camera -z maps to world +x, camera +y to world +y
    iv = np.eye(4)
    iv[:3, 0] = [0, 0, -1]          # camera x axis in world
    iv[:3, 1] = [0, 1, 0]           # camera y axis
    iv[:3, 2] = [-1, 0, 0]          # camera z axis (the camera looks along -z)
    iv[:3, 3] = [3, -2, 7]
    rays = orc.primary_rays(W, H, _colmajor(iv), _colmajor(ip))
    d = rays["dir"].reshape(H, W, 3).astype(np.float64)
    t = math.tan(math.radians(fovy) / 2)
    xs, ys = np.meshgrid(np.arange(W) / W * 2 - 1, np.arange(H) / H * 2 - 1)
    cam = np.stack([xs * t * aspect, ys * t, -np.ones_like(xs)], axis=-1)
    ref = _norm(cam @ iv[:3, :3].T)
    assert np.abs(d - ref).max() < 1e-6
    assert np.abs(d[H // 2, W // 2] - [1, 0, 0]).max() < 1e-7
    assert (rays["org"] == np.float32([3, -2, 7])).all()
