"""Seeded procedural scenes, cameras and ray sets for the five workloads.

Scenes are lists of meshes ``(vertices, indices, global_mesh_number)`` in the
reference's input layout (Vertex: vec4 position with w = 1, half-packed
normal/tangent and UV, ModelFileLoader.cpp:108-155).  All randomness comes from
numpy's PCG64 with fixed seeds; per-pixel ray directions come from the 32-bit
PCG hash (Jarzynski & Olano, "Hash Functions for GPU Rendering", JCGT 2020).

Workloads (sizes are the defaults):
  boxes_scene       16x16 unit-footprint boxes (12 tris each, heights U[0.5,4],
                    seed 1) + 2-tri ground plane: 3,074 triangles
  icosphere_cloud   N icospheres (subdivision s), centres U[0,100]^3, radii
                    lognormal(0, 0.5): 2,000 x 1,280 = 2.56M triangles at s=3
"""
from __future__ import annotations

import math

import numpy as np

from ._lib import RAY_DTYPE, VERTEX_DTYPE

# --- vertex packing --------------------------------------------------------------


def _pack_half2(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    ha = np.asarray(a, np.float32).astype(np.float16).view(np.uint16).astype(np.uint32)
    hb = np.asarray(b, np.float32).astype(np.float16).view(np.uint16).astype(np.uint32)
    return ha | (hb << np.uint32(16))


def make_vertices(pos: np.ndarray, normal: np.ndarray | None = None, uv: np.ndarray | None = None) -> np.ndarray:
    """Vertex records (Vertex.h:7-12) with position.w = 1 and half-packed normal / UV."""
    n = pos.shape[0]
    v = np.zeros(n, dtype=VERTEX_DTYPE)
    v["position"][:, :3] = pos.astype(np.float32)
    v["position"][:, 3] = 1.0
    if normal is not None:
        v["normal_tangent_data"][:, 0] = _pack_half2(normal[:, 0], normal[:, 1])
        v["normal_tangent_data"][:, 1] = _pack_half2(normal[:, 2], np.zeros(n, np.float32))
    if uv is not None:
        v["texcoords"] = _pack_half2(uv[:, 0], uv[:, 1])
    else:
        v["texcoords"] = _pack_half2(np.zeros(n, np.float32), np.zeros(n, np.float32))
    return v


# --- geometry ----------------------------------------------------------------------

_BOX_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                         [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float32)
_BOX_TRIS = np.array([[0, 2, 1], [0, 3, 2],   # z = 0
                      [4, 5, 6], [4, 6, 7],   # z = 1
                      [0, 1, 5], [0, 5, 4],   # y = 0
                      [3, 7, 6], [3, 6, 2],   # y = 1
                      [0, 4, 7], [0, 7, 3],   # x = 0
                      [1, 2, 6], [1, 6, 5]],  # x = 1
                     np.uint32)


def boxes_scene(grid: int = 16, seed: int = 1, hmin: float = 0.5, hmax: float = 4.0, ground: float = 16.0):
    """Config-1 scene: grid x grid boxes on a 1.0 grid + a 2-triangle ground plane.

    Returns one Object's meshes: mesh 0 = all boxes (12 tris each), mesh 1 = ground.
    """
    rng = np.random.default_rng(seed)
    heights = rng.uniform(hmin, hmax, size=grid * grid).astype(np.float32)
    pos, idx = [], []
    for k in range(grid * grid):
        i, j = divmod(k, grid)
        c = _BOX_CORNERS.copy()
        c[:, 0] += i
        c[:, 2] += j
        c[:, 1] *= heights[k]
        pos.append(c)
        idx.append(_BOX_TRIS + 8 * k)
    pos = np.concatenate(pos)
    idx = np.concatenate(idx).ravel()
    lo, hi = -ground, grid + ground
    gpos = np.array([[lo, 0, lo], [hi, 0, lo], [hi, 0, hi], [lo, 0, hi]], np.float32)
    gidx = np.array([0, 2, 1, 0, 3, 2], np.uint32)
    up = np.tile(np.array([[0, 1, 0]], np.float32), (4, 1))
    return [(make_vertices(pos), idx.astype(np.uint32), 0), (make_vertices(gpos, up), gidx, 1)]


def icosphere(subdivisions: int):
    """Unit icosphere: (positions float64 [V,3], triangles uint32 [F,3]); 10*4^s+2 verts, 20*4^s tris."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = [np.array(p, np.float64) / np.linalg.norm(p) for p in verts]
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[key] = len(v) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(v), np.array(faces, np.uint32)


def icosphere_cloud(count: int = 2000, subdivisions: int = 3, seed: int = 2, box: float = 100.0,
                    sigma: float = 0.5, single_mesh: bool = False):
    """Config-2 scene: `count` icospheres, centres U[0, box]^3, radii lognormal(0, sigma).

    One mesh per sphere (global mesh number = sphere index) unless single_mesh.
    """
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, box, size=(count, 3))
    radii = rng.lognormal(0.0, sigma, size=count)
    unit, faces = icosphere(subdivisions)
    nv = unit.shape[0]
    pos = (unit[None, :, :] * radii[:, None, None] + centres[:, None, :]).astype(np.float32)
    normals = np.broadcast_to(unit[None].astype(np.float32), pos.shape)
    if single_mesh:
        idx = (faces[None, :, :] + (np.arange(count, dtype=np.uint32) * nv)[:, None, None]).ravel()
        return [(make_vertices(pos.reshape(-1, 3), normals.reshape(-1, 3)), idx.astype(np.uint32), 0)]
    flat_faces = faces.ravel()
    return [(make_vertices(pos[k], normals[k]), flat_faces, k) for k in range(count)]


def _lattice(ix: np.ndarray, iz: np.ndarray, seed: int) -> np.ndarray:
    """Value-noise lattice values in [0, 1): a PCG hash of (x, z, seed)."""
    h = (ix.astype(np.uint32) * np.uint32(73856093)) ^ (iz.astype(np.uint32) * np.uint32(19349663)) ^ \
        np.uint32((seed * 83492791) & 0xFFFFFFFF)
    return pcg_hash(h).astype(np.float64) / 4294967296.0


def value_noise(x: np.ndarray, z: np.ndarray, seed: int) -> np.ndarray:
    """Smoothly interpolated lattice noise in [0, 1) (smoothstep weights)."""
    x0, z0 = np.floor(x), np.floor(z)
    fx, fz = x - x0, z - z0
    wx, wz = fx * fx * (3 - 2 * fx), fz * fz * (3 - 2 * fz)
    ix, iz = x0.astype(np.int64), z0.astype(np.int64)
    a = _lattice(ix, iz, seed)
    b = _lattice(ix + 1, iz, seed)
    c = _lattice(ix, iz + 1, seed)
    d = _lattice(ix + 1, iz + 1, seed)
    return (a * (1 - wx) + b * wx) * (1 - wz) + (c * (1 - wx) + d * wx) * wz


def fbm(x: np.ndarray, z: np.ndarray, octaves: int = 6, seed: int = 5) -> np.ndarray:
    """Fractal Brownian motion: sum over octaves of 0.5^o * value_noise(2^o * p), normalised to [0, 1)."""
    total = np.zeros_like(x, dtype=np.float64)
    norm = 0.0
    for o in range(octaves):
        amp, freq = 0.5 ** o, 2.0 ** o
        total += amp * value_noise(x * freq, z * freq, seed + 101 * o)
        norm += amp
    return total / norm


def terrain(n: int = 2048, size: float = 2048.0, height: float = 240.0, octaves: int = 6, seed: int = 5,
            features: float = 8.0, mesh: int = 0):
    """Config-5 terrain: an n x n grid of quads (2 n^2 triangles) over [0, size]^2 in x, z, displaced in y
    by height * fBm(6 octaves of seed-driven value noise, `features` base cells across).  Normals from
    central differences of the height field, UV = grid position / n."""
    g = np.arange(n + 1, dtype=np.float64)
    gx, gz = np.meshgrid(g, g)                         # gz rows, gx columns
    hgt = height * fbm(gx * (features / n), gz * (features / n), octaves, seed)
    step = size / n
    pos = np.stack([gx * step, hgt, gz * step], axis=-1).reshape(-1, 3).astype(np.float32)
    dx = np.gradient(hgt, axis=1) / step
    dz = np.gradient(hgt, axis=0) / step
    nrm = np.stack([-dx, np.ones_like(dx), -dz], axis=-1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    uv = np.stack([gx / n, gz / n], axis=-1).reshape(-1, 2)
    v = make_vertices(pos, nrm.reshape(-1, 3).astype(np.float32), uv.astype(np.float32))
    i = np.arange(n, dtype=np.uint32)
    qi, qj = np.meshgrid(i, i)                         # qj rows (z), qi columns (x)
    a = (qj * (n + 1) + qi).ravel()
    b, c, d = a + 1, a + (n + 1), a + (n + 2)
    idx = np.stack([a, c, b, b, c, d], axis=-1).reshape(-1).astype(np.uint32)
    return (v, idx, mesh), hgt


def large_scene(n: int = 2048, spheres: int = 1400, subdivisions: int = 4, seed: int = 5, size: float = 2048.0,
                height: float = 240.0):
    """Config-5 scene: the terrain (mesh 0) plus `spheres` icospheres (meshes 1..spheres) with radii
    lognormal(ln 6, 0.5) resting 0..40 units above the terrain at uniform x, z.  2 n^2 + spheres * 20 * 4^s
    triangles: 8,388,608 + 7,168,000 = 15,556,608 at the defaults."""
    rng = np.random.default_rng(seed)
    ter, hgt = terrain(n, size, height, seed=seed)
    unit, faces = icosphere(subdivisions)
    xs = rng.uniform(0.0, size, spheres)
    zs = rng.uniform(0.0, size, spheres)
    radii = rng.lognormal(math.log(6.0), 0.5, spheres)
    gi = np.clip((xs / size * n).astype(np.int64), 0, n)
    gj = np.clip((zs / size * n).astype(np.int64), 0, n)
    ys = hgt[gj, gi] + radii + rng.uniform(0.0, 40.0, spheres)
    centres = np.stack([xs, ys, zs], axis=-1)
    meshes = [ter]
    flat = faces.ravel()
    nrm = unit.astype(np.float32)
    for k in range(spheres):
        pos = (unit * radii[k] + centres[k]).astype(np.float32)
        meshes.append((make_vertices(pos, nrm), flat, 1 + k))
    return meshes


LARGE_CAMERA = dict(eye=(1024.0, 420.0, -260.0), center=(1024.0, 60.0, 900.0), fovy_deg=60.0)


def scene_arrays(meshes):
    """Concatenated (positions [V,3] f32, triangles [N,3] u32) of a mesh list."""
    pos, tris, off = [], [], 0
    for v, idx, _ in meshes:
        pos.append(v["position"][:, :3])
        tris.append(np.asarray(idx, np.uint32).reshape(-1, 3) + off)
        off += v.shape[0]
    return np.concatenate(pos), np.concatenate(tris)


def concat_meshes(meshes):
    """(vertex records [V], indices [3N] u32 with vertex offsets, mesh id per triangle [N]) in mesh order,
    as BuildBVH concatenates an object's meshes (BVHConstructor.cpp:981-1002)."""
    vs, idx, mids, off = [], [], [], 0
    for v, i, m in meshes:
        vs.append(v)
        idx.append(np.asarray(i, np.uint32).reshape(-1) + np.uint32(off))
        mids.append(np.full(len(i) // 3, m, np.int32))
        off += len(v)
    return np.concatenate(vs), np.concatenate(idx), np.concatenate(mids)


# --- cameras (glm::lookAt / glm::perspective, right-handed, clip z in [-1, 1]) -----


def look_at(eye, center, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    eye, center, up = (np.asarray(x, np.float64) for x in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[0, 3], m[1, 3], m[2, 3] = -s @ eye, -u @ eye, f @ eye
    return m


def perspective(fovy_deg: float, aspect: float, near: float, far: float) -> np.ndarray:
    f = 1.0 / math.tan(math.radians(fovy_deg) / 2.0)
    m = np.zeros((4, 4))
    m[0, 0] = f / aspect
    m[1, 1] = f
    m[2, 2] = (far + near) / (near - far)
    m[2, 3] = 2.0 * far * near / (near - far)
    m[3, 2] = -1.0
    return m


def camera_inverses(eye, center, fovy_deg: float, aspect: float, near: float = 0.1, far: float = 1000.0):
    """(inv_view, inv_proj) as float32 4x4 matrices (math convention)."""
    iv = np.linalg.inv(look_at(eye, center)).astype(np.float32)
    ip = np.linalg.inv(perspective(fovy_deg, aspect, near, far)).astype(np.float32)
    return iv, ip


BOXES_CAMERA = dict(eye=(-8.0, 6.0, -8.0), center=(8.0, 0.0, 8.0), fovy_deg=60.0)
CLOUD_CAMERA = dict(eye=(50.0, 50.0, 2.0), center=(50.0, 50.0, 50.0), fovy_deg=60.0)

# --- per-pixel hashing and secondary rays --------------------------------------------


def pcg_hash(v: np.ndarray) -> np.ndarray:
    """32-bit PCG hash (RXS-M-XS output of one LCG step), vectorised over uint32."""
    v = np.asarray(v, dtype=np.uint32)
    with np.errstate(over="ignore"):
        state = v * np.uint32(747796405) + np.uint32(2891336453)
        word = ((state >> ((state >> np.uint32(28)) + np.uint32(4))) ^ state) * np.uint32(277803737)
    return (word >> np.uint32(22)) ^ word


def hash_uniforms(pixel: np.ndarray, seed: int, k: int) -> np.ndarray:
    """k uniforms in [0, 1) per pixel: chained PCG hashes of pixel ^ seed-mix."""
    h = pcg_hash(np.asarray(pixel, np.uint32) ^ np.uint32((seed * 0x9E3779B9) & 0xFFFFFFFF))
    out = []
    for _ in range(k):
        out.append((h >> np.uint32(8)).astype(np.float32) * np.float32(1.0 / (1 << 24)))
        h = pcg_hash(h)
    return np.stack(out, axis=-1)


def cos_weighted_hemisphere(n: np.ndarray, r: np.ndarray) -> np.ndarray:
    """CosWeightedHemisphere (Shaders/Include/Sampling.glsl:1-12), float32."""
    n = n.astype(np.float32)
    a = np.array([0.0, 1.0, 1.0], np.float32)
    uu = np.cross(n, a)
    ln = np.linalg.norm(uu, axis=1, keepdims=True)
    bad = ln[:, 0] < 1e-6
    if bad.any():  # n parallel to (0,1,1): any perpendicular basis will do
        uu[bad] = np.cross(n[bad], np.array([1.0, 0.0, 0.0], np.float32))
        ln[bad] = np.linalg.norm(uu[bad], axis=1, keepdims=True)
    uu = uu / ln
    vv = np.cross(uu, n)
    two_pi = np.float32(2.0 * 3.14159265359)
    ra = np.sqrt(r[:, 1])
    rx = ra * np.cos(two_pi * r[:, 0])
    ry = ra * np.sin(two_pi * r[:, 0])
    rz = np.sqrt(np.float32(1.0) - r[:, 1])
    rr = rx[:, None] * uu + ry[:, None] * vv + rz[:, None] * n
    return (rr / np.linalg.norm(rr, axis=1, keepdims=True)).astype(np.float32)


def secondary_rays(primary_rays: np.ndarray, primary_hits: np.ndarray, positions: np.ndarray,
                   tris_flat: np.ndarray, seed: int = 2, offset: float = 1e-4, ao_tmax: float = 2.0,
                   gi_tmax: float = 1e6):
    """RTAO any-hit and GI closest-hit rays for every pixel whose primary ray hit.

    G-buffer from the primary hit record: P = o + d*t, N = face normal of the hit
    triangle facing the ray; origin P + N*offset; directions cosine-weighted
    around N from per-pixel PCG uniforms (AO: 1st/2nd, GI: 3rd/4th).
    `tris_flat` is the flattened triangle array (vertex indices in packed[:, :3]).
    Returns (ao_rays, gi_rays) as RAY_DTYPE arrays with id = pixel index.
    """
    hit = primary_hits["t"] > 0
    pix = np.nonzero(hit)[0].astype(np.uint32)
    r = primary_rays[pix]
    h = primary_hits[pix]
    o = r["org"].astype(np.float32)
    d = r["dir"].astype(np.float32)
    p = o + d * h["t"][:, None]
    tv = tris_flat["packed"][h["tri"], :3]
    a, b, c = positions[tv[:, 0]], positions[tv[:, 1]], positions[tv[:, 2]]
    nrm = np.cross(b - a, c - a)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    flip = np.einsum("ij,ij->i", nrm, d) > 0
    nrm[flip] *= -1.0
    nrm = nrm.astype(np.float32)
    origin = (p + nrm * np.float32(offset)).astype(np.float32)
    u = hash_uniforms(pix, seed, 4)
    out = []
    for k, tmax in ((0, ao_tmax), (2, gi_tmax)):
        rays = np.zeros(pix.size, RAY_DTYPE)
        rays["org"] = origin
        rays["dir"] = cos_weighted_hemisphere(nrm, u[:, k:k + 2])
        rays["tmax"] = np.float32(tmax)
        rays["id"] = pix
        out.append(rays)
    return out[0], out[1]
