"""Seeded workloads built on the library: scenes, BVHs on the device and ray batches.

incoherent_cloud()  2,000 icospheres (2.56M triangles); a 1920x1080 G-buffer from
                    primary closest hits; per hit pixel one RTAO any-hit ray (tmax 2)
                    and one GI closest-hit ray.
batched_scenes()    N independent icosphere clouds with triangle counts log-uniform
                    in [1k, 2M], each with its own BVH; per scene a ray batch of size
                    log-uniform in [4k, 2M] split 40/40/20 into coherent primary rays,
                    cosine GI rays and any-hit shadow rays toward a random point light.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import scenes
from ._lib import HIT_DTYPE, QUERY_ANY, QUERY_CLOSEST, RAY_DTYPE, FORMAT_STACKLESS


class DeviceSpan:
    """A sub-range of a DeviceArray (records [offset, offset+count))."""

    def __init__(self, arr, offset: int, count: int):
        assert 0 <= offset and offset + count <= arr.count
        self.base = arr
        self.ptr = arr.ptr + offset * arr.dtype.itemsize
        self.count = int(count)
        self.dtype = arr.dtype


@dataclass
class CloudWorkload:
    intersector: object
    scene: object
    triangles: int
    ao: np.ndarray
    gi: np.ndarray
    host: dict
    build_s: float
    primary_hits: int


def incoherent_cloud(crt, spheres: int = 2000, width: int = 1920, height: int = 1080, fmt: int = FORMAT_STACKLESS,
                     seed: int = 2) -> CloudWorkload:
    import time
    meshes = scenes.icosphere_cloud(count=spheres, subdivisions=3, seed=seed)
    r = crt.RayIntersector(fmt)
    t0 = time.perf_counter()
    r.add_object(0, meshes)
    build_s = time.perf_counter() - t0
    r.push_entity(0)
    r.buffer_entities()
    r.buffer_data()
    hb = r.host_buffers()
    iv, ip = scenes.camera_inverses(aspect=width / height, **scenes.CLOUD_CAMERA)
    drays, dhits = r.intersect_primary(width, height, iv, ip)
    crt.lib().crt_device_sync()
    prays, phits = drays.download(), dhits.download()
    ao, gi = scenes.secondary_rays(prays, phits, hb["vertices"]["position"][:, :3], hb["triangles"], seed=seed)
    return CloudWorkload(r, r.scene(), int(hb["triangles"].size), ao, gi, hb, build_s, int(ao.size))


@dataclass
class SceneBatch:
    index: int
    triangles: int
    intersector: object
    scene: object
    host: dict
    primary: int          # rays [0, primary) of `closest` are primary rays
    gi: int               # rays [primary, primary+gi) of `closest` are GI rays
    closest: np.ndarray   # primary + GI rays (closest hit)
    shadow: np.ndarray    # shadow rays (any hit)
    d_closest: object = None
    d_shadow: object = None
    h_closest: object = None
    h_shadow: object = None
    wide: dict = field(default_factory=dict)   # (width, max_leaf) -> DeviceWide

    def scene_for(self, layout):
        """The BVH2 scene (layout None) or the wide layout (width, max_leaf)."""
        return self.scene if layout is None else self.wide[layout].scene


@dataclass
class BatchedWorkload:
    scenes: list = field(default_factory=list)
    build_s: float = 0.0

    @property
    def rays(self) -> int:
        return sum(s.closest.size + s.shadow.size for s in self.scenes)

    @property
    def triangles(self) -> int:
        return sum(s.triangles for s in self.scenes)

    def batches(self, crt, layout=None):
        """(scene, query, rays, hits) per ray type: primary, GI (closest) and shadow (any) of every scene,
        over the BVH2 (layout None) or the wide layout (width, max_leaf) built by build_wide."""
        out = []
        for s in self.scenes:
            sc = s.scene_for(layout)
            out.append((sc, QUERY_CLOSEST, DeviceSpan(s.d_closest, 0, s.primary),
                        DeviceSpan(s.h_closest, 0, s.primary)))
            out.append((sc, QUERY_CLOSEST, DeviceSpan(s.d_closest, s.primary, s.gi),
                        DeviceSpan(s.h_closest, s.primary, s.gi)))
            out.append((sc, QUERY_ANY, s.d_shadow, s.h_shadow))
        return out

    def build_wide(self, crt, width: int, max_leaf: int = 2, collapse: int = 0, tri_cost: float = 0.0) -> float:
        """Derives and uploads the wide layout of every scene (key (width, max_leaf), with the collapse as a
        third element when it is not the greedy one); returns the host seconds spent."""
        import time
        t0 = time.perf_counter()
        key = (width, max_leaf) if not collapse else (width, max_leaf, collapse)
        for s in self.scenes:
            if key not in s.wide:
                s.wide[key] = crt.upload_wide(FORMAT_STACKLESS, s.host, width, max_leaf, s.scene.triangles,
                                              s.scene.vertices, collapse, 0.0, tri_cost)
        return time.perf_counter() - t0


def _log_uniform(rng, lo: float, hi: float, n: int) -> np.ndarray:
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size=n))


def batched_scenes(crt, count: int = 64, seed: int = 3, tri_range=(1e3, 2e6), ray_range=(4e3, 2e6),
                   fmt: int = FORMAT_STACKLESS) -> BatchedWorkload:
    """Builds the scenes and their BVHs (CPU constructor), uploads them and generates every ray batch.

    Each cloud keeps the 2,000-sphere cloud's density: its box side is 100 * (spheres / 2000)^(1/3).
    The camera sits in front of the box looking at its centre (60 deg FOV, 16:9).
    """
    import time
    rng = np.random.default_rng(seed)
    tris = _log_uniform(rng, *tri_range, count)
    nrays = _log_uniform(rng, *ray_range, count).astype(np.int64)
    lights_u = rng.uniform(0.0, 1.0, size=(count, 3))
    work = BatchedWorkload()
    for k in range(count):
        spheres = max(1, int(round(tris[k] / 1280.0)))
        box = 100.0 * (spheres / 2000.0) ** (1.0 / 3.0)
        meshes = scenes.icosphere_cloud(count=spheres, subdivisions=3, seed=1000 + k, box=box)
        r = crt.RayIntersector(fmt)
        t0 = time.perf_counter()
        r.add_object(0, meshes)
        work.build_s += time.perf_counter() - t0
        r.push_entity(0)
        r.buffer_entities()
        r.buffer_data()
        hb = r.host_buffers()
        n = int(nrays[k])
        n_p = int(round(0.4 * n))
        n_gi = int(round(0.4 * n))
        n_sh = n - n_p - n_gi
        # coherent primary rays: the first n_p pixels (row-major) of a 16:9 image
        h = max(1, int(math.ceil(math.sqrt(n_p * 9.0 / 16.0))))
        w = max(1, int(math.ceil(n_p / h)))
        cam = dict(eye=(box / 2, box / 2, -0.35 * box), center=(box / 2, box / 2, box / 2), fovy_deg=60.0)
        iv, ip = scenes.camera_inverses(aspect=w / h, **cam)
        drays, dhits = r.intersect_primary(w, h, iv, ip)
        crt.lib().crt_device_sync()
        prays, phits = drays.download()[:n_p], dhits.download()[:n_p]
        del drays, dhits
        pos = hb["vertices"]["position"][:, :3]
        gi = _surface_rays(prays, phits, pos, hb["triangles"], n_gi, seed=seed * 7919 + k, box=box, kind="gi")
        light = lights_u[k] * box
        sh = _surface_rays(prays, phits, pos, hb["triangles"], n_sh, seed=seed * 104729 + k, box=box, kind="shadow",
                           light=light)
        closest = np.concatenate([prays, gi])
        closest["id"] = np.arange(closest.size, dtype=np.uint32)
        sb = SceneBatch(k, int(hb["triangles"].size), r, r.scene(), hb, n_p, n_gi, closest, sh)
        sb.d_closest = crt.DeviceArray.from_host(closest)
        sb.d_shadow = crt.DeviceArray.from_host(sh)
        sb.h_closest = crt.DeviceArray(HIT_DTYPE, closest.size)
        sb.h_shadow = crt.DeviceArray(HIT_DTYPE, sh.size)
        work.scenes.append(sb)
    return work


def _surface_rays(prays, phits, positions, tris_flat, n: int, seed: int, box: float, kind: str, light=None):
    """n rays leaving the primary hit points (cycled): cosine GI directions, or toward `light` (shadow).

    A scene whose primary rays all missed gets rays from uniform points of its box instead."""
    rng = np.random.default_rng(seed)
    out = np.zeros(n, RAY_DTYPE)
    if n == 0:
        return out
    hit = np.nonzero(phits["t"] > 0)[0]
    if hit.size == 0:
        o = rng.uniform(0, box, size=(n, 3)).astype(np.float32)
        if kind == "shadow":
            d = light[None, :].astype(np.float64) - o
            dist = np.linalg.norm(d, axis=1)
            out["dir"] = (d / np.maximum(dist[:, None], 1e-12)).astype(np.float32)
            out["tmax"] = dist.astype(np.float32)
        else:
            d = rng.normal(size=(n, 3))
            out["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
            out["tmax"] = np.float32(1e6)
        out["org"] = o
        out["id"] = np.arange(n, dtype=np.uint32)
        return out
    pick = hit[np.arange(n) % hit.size]
    r = prays[pick]
    h = phits[pick]
    p = r["org"] + r["dir"] * h["t"][:, None]
    tv = tris_flat["packed"][h["tri"], :3]
    a, b, c = positions[tv[:, 0]], positions[tv[:, 1]], positions[tv[:, 2]]
    nrm = np.cross(b - a, c - a)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    flip = np.einsum("ij,ij->i", nrm, r["dir"]) > 0
    nrm[flip] *= -1.0
    nrm = nrm.astype(np.float32)
    origin = (p + nrm * np.float32(1e-4)).astype(np.float32)
    out["org"] = origin
    if kind == "shadow":
        d = light[None, :].astype(np.float64) - origin
        dist = np.linalg.norm(d, axis=1)
        out["dir"] = (d / np.maximum(dist[:, None], 1e-12)).astype(np.float32)
        out["tmax"] = dist.astype(np.float32)
    else:
        u = rng.uniform(0, 1, size=(n, 2)).astype(np.float32)
        out["dir"] = scenes.cos_weighted_hemisphere(nrm, u)
        out["tmax"] = np.float32(1e6)
    out["id"] = np.arange(n, dtype=np.uint32)
    return out


@dataclass
class ShadingProblem:
    intersector: object
    scene: object
    host: dict
    rays: object          # primary rays (device)
    hits: object          # primary hits (device)
    model: object         # ShadingModel with the parameters theta
    albedo: np.ndarray
    roughness: np.ndarray
    light_pos: np.ndarray
    intensity: np.ndarray
    target_albedo: np.ndarray
    target_roughness: np.ndarray
    target_intensity: np.ndarray


def shading_problem(crt, size: int = 1024, seed: int = 4, nlights: int = 8, fmt: int = FORMAT_STACKLESS):
    """The 16x16 boxes scene at size x size with per-triangle albedo U[0.05,0.95]^3 and roughness
    U[0.05,1], `nlights` point lights (intensity U[1,10]) above the grid; the target parameters are
    theta + N(0, 0.1) (roughness kept >= 0.02)."""
    rng = np.random.default_rng(seed)
    r = crt.RayIntersector(fmt)
    r.add_object(0, scenes.boxes_scene())
    r.push_entity(0)
    r.buffer_entities()
    r.buffer_data()
    hb = r.host_buffers()
    ntri = int(hb["triangles"].size)
    iv, ip = scenes.camera_inverses(aspect=1.0, **scenes.BOXES_CAMERA)
    drays, dhits = r.intersect_primary(size, size, iv, ip)
    albedo = rng.uniform(0.05, 0.95, size=(ntri, 3))
    rough = rng.uniform(0.05, 1.0, size=ntri)
    lpos = rng.uniform([-4.0, 6.0, -4.0], [20.0, 14.0, 20.0], size=(nlights, 3))
    inten = rng.uniform(1.0, 10.0, size=nlights)
    t_alb = albedo + rng.normal(0.0, 0.1, size=albedo.shape)
    t_rough = np.maximum(rough + rng.normal(0.0, 0.1, size=rough.shape), 0.02)
    t_inten = inten + rng.normal(0.0, 0.1, size=inten.shape)
    model = crt.ShadingModel(albedo, rough, lpos, inten)
    return ShadingProblem(r, r.scene(), hb, drays, dhits, model, albedo, rough, lpos, inten, t_alb, t_rough, t_inten)
