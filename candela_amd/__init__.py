"""candela_amd -- MI355X-native ray traversal for Candela.

Python face of libcandela_rt (include/candela_rt.h).  It mirrors the
reference's host API, ``RayIntersector<T>`` (Source/Core/BVH/Intersector.h:60-124),
and its CPU constructor, ``BuildBVH`` (Source/Core/BVH/BVHConstructor.cpp:951-1108).
Every call goes to the compiled library; there is no Python fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from ._lib import (KERNEL_AUTO, KERNEL_PERSISTENT, KERNEL_THREAD_PER_RAY, KERNEL_WAVE, KERNEL_SPECULATIVE,
                   KERNEL_SPECULATIVE_REFILL, BOUNDS_DTYPE, ENTITY_DTYPE, FLAG_IGNORE_TRANSPARENT, FLAG_TRI0_IS_MISS, FLAG_SLOT_ORDER, FORMAT_STACK,
                   FORMAT_STACKLESS, FORMAT_WIDE4, FORMAT_WIDE8, WIDE_GROUP_DTYPE, TRI_RECORD_DTYPE, WideStats, WideOptions,
                   COLLAPSE_GREEDY, COLLAPSE_SAH,
                   TEXTURE_REF_DTYPE, TEXTURE_DTYPE, HIT_ATTR_DTYPE, PATH_PARAMS_DTYPE, PATH_SAMPLE_DTYPE,
                   HIT_DTYPE, QUERY_ANY, QUERY_CLOSEST, RAY_DTYPE, STACK_NODE_DTYPE,
                   STACKLESS_NODE_DTYPE, STATUS_BAD_INDEX, STATUS_ITER_CAP, STATUS_STACK_OVERFLOW,
                   STATUS_VISITS_MASK, SWAP_NONE, SWAP_SEEDED, TRIANGLE_DTYPE, VERTEX_DTYPE, BuildOptions,
                   BvhStats, CrtError, Mesh, SceneView, TraceConfig, Batch, BatchConfig, ShadeParams,
                   SHADE_RECORD_DTYPE, RAYGEN_PARAMS_DTYPE, PROBE_GRID_DTYPE, CacheHeader, E_CORRUPT, E_IO, E_STALE, WIDE_GPU_BAD_LINK, WIDE_GPU_NOT_TREE,
                   WIDE_GPU_BAD_VERTEX, WIDE_GPU_TOO_DEEP, WIDE_GPU_BAD_LEAF, WIDE_GPU_MAX_LEVELS, QUEUE_BYTES, check, lib, node_dtype, ptr)

__all__ = [
    "VERTEX_DTYPE", "TRIANGLE_DTYPE", "STACKLESS_NODE_DTYPE", "STACK_NODE_DTYPE", "BOUNDS_DTYPE", "ENTITY_DTYPE",
    "RAY_DTYPE", "HIT_DTYPE", "FORMAT_STACKLESS", "FORMAT_STACK", "QUERY_CLOSEST", "QUERY_ANY",
    "FLAG_IGNORE_TRANSPARENT", "FLAG_TRI0_IS_MISS", "FLAG_SLOT_ORDER", "STATUS_ITER_CAP", "STATUS_STACK_OVERFLOW", "STATUS_BAD_INDEX",
    "STATUS_VISITS_MASK", "SWAP_NONE", "SWAP_SEEDED", "KERNEL_AUTO", "KERNEL_THREAD_PER_RAY", "KERNEL_PERSISTENT",
    "KERNEL_WAVE", "CrtError", "DeviceArray", "Stream", "BuiltBVH", "build_bvh",
    "RayIntersector", "trace", "device_count", "node_dtype", "lib", "Batch", "BatchConfig", "BatchPlan",
    "GpuWide",
    "autotune_batched", "ShadeParams", "SHADE_RECORD_DTYPE", "ShadingModel", "FORMAT_WIDE4", "FORMAT_WIDE8",
    "WIDE_GROUP_DTYPE", "TRI_RECORD_DTYPE", "WideBVH", "wide_build", "DeviceWide", "upload_wide",
    "COLLAPSE_GREEDY", "COLLAPSE_SAH", "WideOptions",
    "TEXTURE_REF_DTYPE", "TEXTURE_DTYPE", "HIT_ATTR_DTYPE", "MaterialTable", "hit_attributes",
    "PATH_PARAMS_DTYPE", "PATH_SAMPLE_DTYPE", "path_params", "trace_paths", "hash64", "bvh_key", "load_bvh",
    "CacheHeader", "E_IO", "E_CORRUPT", "E_STALE", "RAYGEN_PARAMS_DTYPE", "PROBE_GRID_DTYPE", "raygen_params",
    "RayGen", "probe_grid", "gen_probe_rays", "GpuBVH", "Graph", "WIDE_GPU_BAD_LINK", "WIDE_GPU_NOT_TREE",
    "WIDE_GPU_BAD_VERTEX", "WIDE_GPU_TOO_DEEP", "WIDE_GPU_BAD_LEAF", "WIDE_GPU_MAX_LEVELS", "QUEUE_BYTES",
    "queue_block", "autotune_batched_layouts", "KERNEL_SPECULATIVE", "KERNEL_SPECULATIVE_REFILL",
]


def device_count() -> int:
    return int(lib().crt_device_count())


class Stream:
    """A non-blocking HIP stream owned by the library."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().crt_stream_create(C.byref(h)))
        self.handle = h.value

    def sync(self) -> None:
        check(lib().crt_stream_sync(self.handle))

    def close(self) -> None:
        if self.handle:
            lib().crt_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """Work captured from a Stream (crt_graph_begin / crt_graph_end) and replayed with launch()."""

    def __init__(self, stream: "Stream"):
        self.stream = stream
        self.handle = None

    def __enter__(self):
        check(lib().crt_graph_begin(self.stream.handle))
        return self

    def __exit__(self, exc_type, exc, tb):
        h = C.c_void_p()
        rc = lib().crt_graph_end(self.stream.handle, C.byref(h))
        if exc_type is None:
            check(rc)
            self.handle = h.value
        elif rc == 0 and h.value:   # the capture failed part way: drop what it recorded
            lib().crt_graph_destroy(h.value)
        return False

    def launch(self, stream: "Stream | None" = None) -> None:
        check(lib().crt_graph_launch(self.handle, _stream_handle(stream or self.stream)))

    def close(self) -> None:
        if self.handle:
            lib().crt_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stream_handle(stream) -> int | None:
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else int(stream)


class DeviceArray:
    """A device allocation holding `count` records of `dtype`."""

    def __init__(self, dtype, count: int):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.dtype.itemsize * self.count
        p = C.c_void_p()
        check(lib().crt_device_alloc(C.byref(p), max(1, self.nbytes)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray, stream=None) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.dtype, a.size)
        d.upload(a, stream)
        return d

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes != self.nbytes:
            raise ValueError(f"upload size {a.nbytes} != allocation {self.nbytes}")
        check(lib().crt_memcpy_h2d(self.ptr, ptr(a), self.nbytes, _stream_handle(stream)))

    def download(self, stream=None) -> np.ndarray:
        out = np.empty(self.count, dtype=self.dtype)
        check(lib().crt_memcpy_d2h(ptr(out), self.ptr, self.nbytes, _stream_handle(stream)))
        return out

    def fill_bytes(self, value: int, stream=None) -> None:
        check(lib().crt_memset_d(self.ptr, int(value), self.nbytes, _stream_handle(stream)))

    def free(self) -> None:
        if self.ptr:
            lib().crt_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _meshes(meshes):
    """[(vertices VERTEX_DTYPE, indices uint32, global_mesh_number)] -> ctypes array + keep-alive list."""
    keep = []
    arr = (Mesh * max(1, len(meshes)))()
    for i, (v, idx, num) in enumerate(meshes):
        v = np.ascontiguousarray(v, dtype=VERTEX_DTYPE)
        idx = np.ascontiguousarray(idx, dtype=np.uint32).ravel()
        keep += [v, idx]
        arr[i] = Mesh(ptr(v), v.size, ptr(idx), idx.size, int(num))
    return arr, keep


@dataclass
class BuiltBVH:
    format: int
    nodes: np.ndarray
    triangles: np.ndarray
    vertices: np.ndarray
    stats: dict


def _bvh_copies(h, fmt: int) -> BuiltBVH:
    st = BvhStats()
    check(lib().crt_bvh_get_stats(h, C.byref(st)))
    n_nodes, n_tri, n_vtx = int(st.array_length), int(st.triangle_count), int(st.vertex_count)
    nd = node_dtype(fmt)
    nodes = np.frombuffer((C.c_char * (n_nodes * nd.itemsize)).from_address(lib().crt_bvh_nodes(h)),
                          dtype=nd).copy() if n_nodes else np.zeros(0, nd)
    tris = np.frombuffer((C.c_char * (n_tri * 16)).from_address(lib().crt_bvh_triangles(h)),
                         dtype=TRIANGLE_DTYPE).copy() if n_tri else np.zeros(0, TRIANGLE_DTYPE)
    verts = np.frombuffer((C.c_char * (n_vtx * 32)).from_address(lib().crt_bvh_vertices(h)),
                          dtype=VERTEX_DTYPE).copy() if n_vtx else np.zeros(0, VERTEX_DTYPE)
    stats = {k: int(getattr(st, k)) for k, _ in BvhStats._fields_}
    return BuiltBVH(fmt, nodes, tris, verts, stats)


def hash64(data, seed: int = 0) -> int:
    """crt_hash64: XXH64 of a bytes-like object or contiguous array."""
    b = np.frombuffer(memoryview(data).cast("B"), np.uint8) if not isinstance(data, np.ndarray) else \
        np.ascontiguousarray(data).view(np.uint8).reshape(-1)
    return int(lib().crt_hash64(ptr(b) if b.size else None, b.size, seed & 0xFFFFFFFFFFFFFFFF))


def bvh_key(fmt: int, meshes, child_swap: int = SWAP_NONE, seed: int = 0) -> int:
    """crt_bvh_key: the cache key of a build's inputs."""
    arr, keep = _meshes(meshes)
    opts = BuildOptions(child_swap, seed & 0xFFFFFFFF, 0)
    k = C.c_uint64()
    check(lib().crt_bvh_key(fmt, arr, len(meshes), C.byref(opts), C.byref(k)))
    del keep
    return int(k.value)


def _header_dict(hd: CacheHeader) -> dict:
    out = {}
    for name, _ in CacheHeader._fields_:
        v = getattr(hd, name)
        if name in ("magic", "toolchain"):
            v = bytes(v).decode()
        elif name.startswith("section_"):
            v = [int(x) for x in v]
        elif name.startswith("reserved"):
            continue
        out[name] = v
    return out


def load_bvh(path: str, expect_key: int = 0, t_offset: int | None = None) -> tuple[BuiltBVH, dict]:
    """crt_bvh_load (+ crt_bvh_rebase when t_offset is given): host copies and the header."""
    h = C.c_void_p()
    hd = CacheHeader()
    check(lib().crt_bvh_load(os.fsencode(path), expect_key & 0xFFFFFFFFFFFFFFFF, C.byref(h), C.byref(hd)))
    try:
        if t_offset is not None:
            check(lib().crt_bvh_rebase(h, int(t_offset)))
        return _bvh_copies(h, int(hd.format)), _header_dict(hd)
    finally:
        lib().crt_bvh_free(h)


def build_bvh(fmt: int, meshes, t_offset: int = 0, child_swap: int = SWAP_NONE, seed: int = 0,
              threads: int = 0, save_to: str | None = None, rebase_to: int | None = None) -> BuiltBVH:
    """BuildBVH (BVHConstructor.cpp:951-1108) on the CPU; returns host copies of the flattened buffers.
    save_to: also write a cache file (crt_bvh_save with the inputs' crt_bvh_key), before any rebase.
    rebase_to: move the leaf starts to this t_offset (crt_bvh_rebase) before copying."""
    arr, keep = _meshes(meshes)
    opts = BuildOptions(child_swap, seed & 0xFFFFFFFF, threads)
    h = C.c_void_p()
    check(lib().crt_build_bvh(fmt, arr, len(meshes), int(t_offset), C.byref(opts), C.byref(h)))
    try:
        if save_to is not None:
            k = C.c_uint64()
            check(lib().crt_bvh_key(fmt, arr, len(meshes), C.byref(opts), C.byref(k)))
            check(lib().crt_bvh_save(h, k.value, os.fsencode(save_to)))
        if rebase_to is not None:
            check(lib().crt_bvh_rebase(h, int(rebase_to)))
        return _bvh_copies(h, fmt)
    finally:
        lib().crt_bvh_free(h)
        del keep


@dataclass
class WideBVH:
    width: int
    groups: np.ndarray     # WIDE_GROUP_DTYPE [node_count * width / 4]
    records: np.ndarray    # TRI_RECORD_DTYPE [triangles]
    entities: np.ndarray   # ENTITY_DTYPE, node_offset / node_count = wide root / node count
    stats: dict

    @property
    def format(self) -> int:
        return FORMAT_WIDE8 if self.width == 8 else FORMAT_WIDE4


def _view(addr: int, count: int, dt: np.dtype) -> np.ndarray:
    if not count:
        return np.zeros(0, dt)
    return np.frombuffer((C.c_char * (count * dt.itemsize)).from_address(addr), dtype=dt).copy()


def wide_build(src_format: int, nodes: np.ndarray, triangles: np.ndarray, vertices: np.ndarray,
               entities: np.ndarray, width: int = 4, max_leaf: int = 2, collapse: int = COLLAPSE_GREEDY,
               node_cost: float = 0.0, tri_cost: float = 0.0) -> WideBVH:
    """crt_wide_build / crt_wide_build_ex: the BVH4 / BVH8 layout derived from BVH2 buffers (host copies).
    collapse: COLLAPSE_GREEDY (crt_wide_build's) or COLLAPSE_SAH with the node / triangle costs (0 = defaults)."""
    nodes = np.ascontiguousarray(nodes, dtype=node_dtype(src_format))
    triangles = np.ascontiguousarray(triangles, dtype=TRIANGLE_DTYPE)
    vertices = np.ascontiguousarray(vertices, dtype=VERTEX_DTYPE)
    entities = np.ascontiguousarray(entities, dtype=ENTITY_DTYPE)
    h = C.c_void_p()
    args = (src_format, ptr(nodes), nodes.size, ptr(triangles), triangles.size, ptr(vertices), vertices.size,
            ptr(entities), entities.size)
    if collapse == COLLAPSE_GREEDY and node_cost == 0.0 and tri_cost == 0.0:
        check(lib().crt_wide_build(*args, int(width), int(max_leaf), C.byref(h)))
    else:
        opt = WideOptions(int(width), int(max_leaf), int(collapse), float(node_cost), float(tri_cost))
        check(lib().crt_wide_build_ex(*args, C.byref(opt), C.byref(h)))
    try:
        st = WideStats()
        check(lib().crt_wide_get_stats(h, C.byref(st)))
        stats = {k: (float(getattr(st, k)) if k == "mean_fill" else int(getattr(st, k))) for k, _ in WideStats._fields_}
        groups = _view(lib().crt_wide_groups(h), stats["node_count"] * width // 4, WIDE_GROUP_DTYPE)
        records = _view(lib().crt_wide_records(h), stats["record_count"], TRI_RECORD_DTYPE)
        ents = _view(lib().crt_wide_entities(h), entities.size, ENTITY_DTYPE)
    finally:
        lib().crt_wide_free(h)
    return WideBVH(int(width), groups, records, ents, stats)


class MaterialTable:
    """Per-mesh material records (crt_texture_ref) and RGBA8 sRGB textures, on the device."""

    def __init__(self, refs: np.ndarray, textures=()):
        """refs: TEXTURE_REF_DTYPE per mesh id; textures: (texels uint32 [h, w], ...) in table order."""
        self.refs = DeviceArray.from_host(np.ascontiguousarray(refs, TEXTURE_REF_DTYPE))
        self.texels = [DeviceArray.from_host(np.ascontiguousarray(t, np.uint32).reshape(-1)) for t in textures]
        desc = np.zeros(max(1, len(self.texels)), TEXTURE_DTYPE)
        for k, (t, d) in enumerate(zip(textures, self.texels)):
            h, w = np.asarray(t).shape
            desc[k] = (d.ptr, w, h)
        self.desc = DeviceArray.from_host(desc)
        self.nrefs = int(np.asarray(refs).size)
        self.ntextures = len(self.texels)


def hit_attributes(scene: SceneView, materials: MaterialTable, hits: DeviceArray, out: DeviceArray,
                   n: int | None = None, stream=None) -> None:
    """crt_hit_attributes: GetData for n closest-hit records (asynchronous)."""
    n = hits.count if n is None else int(n)
    if n > hits.count or n > out.count:
        raise ValueError("hit / attribute buffers smaller than n")
    check(lib().crt_hit_attributes(C.byref(scene), materials.refs.ptr, materials.nrefs, materials.desc.ptr,
                                   materials.ntextures, hits.ptr, n, out.ptr, _stream_handle(stream)))


def path_params(width: int, height: int, spp: int, inv_view, inv_proj, albedo=None, seed: int = 5,
                max_bounces: int = 2, rr: float = 0.8, tmax: float = 1e6, sky=(1.0, 1.0, 1.0)) -> np.ndarray:
    """A crt_path_params record; `albedo` is a device array of nmesh * 3 floats (or None)."""
    p = np.zeros(1, PATH_PARAMS_DTYPE)
    p["inv_view"], p["inv_proj"] = _colmajor(inv_view), _colmajor(inv_proj)
    p["width"], p["height"], p["spp"], p["seed"] = width, height, spp, seed & 0xFFFFFFFF
    p["max_bounces"], p["rr"], p["tmax"], p["sky"] = max_bounces, rr, tmax, sky
    if albedo is not None:
        p["albedo"], p["nmesh"] = albedo.ptr, albedo.count // 3
    return p


def trace_paths(scene: SceneView, params: np.ndarray, out: DeviceArray, npaths: int | None = None, variant: int = 0,
                grid: int = 0, stream=None, queue: "DeviceArray | None" = None) -> None:
    """crt_trace_paths: paths [0, npaths) into out (PATH_SAMPLE_DTYPE), asynchronous.  queue: a work-queue block
    (queue_block()) the launch uses instead of one leased from the library's pool (required under capture)."""
    params = np.ascontiguousarray(params, PATH_PARAMS_DTYPE).reshape(1)
    npaths = out.count if npaths is None else int(npaths)
    if npaths > out.count:
        raise ValueError("output smaller than npaths")
    cfg = TraceConfig(0, 0, 0, variant, grid, 0, None, queue.ptr if queue is not None else None)
    check(lib().crt_trace_paths(C.byref(scene), ptr(params), out.ptr, npaths, C.byref(cfg), _stream_handle(stream)))


def raygen_params(seed: int = 2, offset: float = 1e-4, ao_tmax: float = 2.0, gi_tmax: float = 1e6,
                  spec_offset=(0.05, 0.1), spec_tmax: float = 1e6, roughness_scale: float = 0.9,
                  roughness=None) -> np.ndarray:
    """A crt_raygen_params record; `roughness` is a device array of one float per mesh id (or None)."""
    p = np.zeros(1, RAYGEN_PARAMS_DTYPE)
    p["seed"], p["offset"], p["ao_tmax"], p["gi_tmax"] = seed & 0xFFFFFFFF, offset, ao_tmax, gi_tmax
    p["spec_offset"], p["spec_tmax"], p["roughness_scale"] = spec_offset, spec_tmax, roughness_scale
    if roughness is not None:
        p["roughness"], p["nmesh"] = roughness.ptr, roughness.count
    return p


class RayGen:
    """Wavefront ray generation (crt_gen_diffuse_rays / crt_gen_specular_rays) with its workspace and
    device ray counter, for inputs of up to `capacity` records."""

    def __init__(self, capacity: int):
        self.capacity = int(capacity)
        self.workspace = DeviceArray(np.uint8, int(lib().crt_raygen_workspace_bytes(self.capacity)))
        self.count = DeviceArray(np.int64, 1)

    def _n(self, hits, n):
        n = hits.count if n is None else int(n)
        if n > self.capacity:
            raise ValueError(f"{n} records exceed the generator's capacity {self.capacity}")
        return n

    def diffuse(self, scene: SceneView, primary: DeviceArray, hits: DeviceArray, params: np.ndarray,
                ao_out: DeviceArray | None, gi_out: DeviceArray | None, n: int | None = None, stream=None) -> None:
        """AO and GI rays of the hit records, compacted in input order (asynchronous; see last_count)."""
        n = self._n(hits, n)
        for o in (ao_out, gi_out):
            if o is not None and o.count < n:
                raise ValueError("output smaller than the input")
        params = np.ascontiguousarray(params, RAYGEN_PARAMS_DTYPE).reshape(1)
        check(lib().crt_gen_diffuse_rays(C.byref(scene), primary.ptr, hits.ptr, n, ptr(params),
                                         ao_out.ptr if ao_out is not None else None,
                                         gi_out.ptr if gi_out is not None else None, self.count.ptr,
                                         self.workspace.ptr, self.workspace.nbytes, _stream_handle(stream)))

    def specular(self, scene: SceneView, primary: DeviceArray, hits: DeviceArray, params: np.ndarray,
                 out: DeviceArray, n: int | None = None, stream=None) -> None:
        """Reflection rays of the hit records, compacted in input order (asynchronous; see last_count)."""
        n = self._n(hits, n)
        if out.count < n:
            raise ValueError("output smaller than the input")
        params = np.ascontiguousarray(params, RAYGEN_PARAMS_DTYPE).reshape(1)
        check(lib().crt_gen_specular_rays(C.byref(scene), primary.ptr, hits.ptr, n, ptr(params), out.ptr,
                                          self.count.ptr, self.workspace.ptr, self.workspace.nbytes,
                                          _stream_handle(stream)))

    def last_count(self, stream=None) -> int:
        """Rays written by the last call (waits for it)."""
        return int(self.count.download(stream)[0])


def probe_grid(res, box_origin, size, seed: int = 0, tmax: float = 1e6) -> np.ndarray:
    g = np.zeros(1, PROBE_GRID_DTYPE)
    g["res"], g["box_origin"], g["size"], g["seed"], g["tmax"] = res, box_origin, size, seed & 0xFFFFFFFF, tmax
    return g


def gen_probe_rays(grid: np.ndarray, out: DeviceArray, stream=None) -> None:
    """crt_gen_probe_rays: one ray per probe into `out` (asynchronous)."""
    grid = np.ascontiguousarray(grid, PROBE_GRID_DTYPE).reshape(1)
    if out.count < int(np.prod(grid["res"][0].astype(np.int64))):
        raise ValueError("output smaller than the probe count")
    check(lib().crt_gen_probe_rays(ptr(grid), out.ptr, _stream_handle(stream)))


def _identity_entity(node_count: int) -> np.ndarray:
    e = np.zeros(1, ENTITY_DTYPE)
    e["model"][0] = np.eye(4, dtype=np.float32).reshape(16)
    e["inverse"][0] = np.eye(4, dtype=np.float32).reshape(16)
    e["node_count"] = node_count
    alpha = np.array([1.0], np.float32).view(np.int32)[0]
    e["data"][0, 1] = alpha
    return e


class GpuBVH:
    """crt_build_bvh_gpu: the stack layout of one object built on the device (a linear BVH over Morton codes;
    not BuildBVH's tree).  `scene` is a crt_scene_view over it with one identity entity."""

    def __init__(self, vertices: np.ndarray, indices: np.ndarray, mesh_ids=None, mesh_id: int = 0, stream=None,
                 t_offset: int = 0):
        v = np.ascontiguousarray(vertices, VERTEX_DTYPE)
        self.t_offset = int(t_offset)
        idx = np.ascontiguousarray(indices, np.uint32).reshape(-1)
        if idx.size % 3 or idx.size < 6 or int(idx.max()) >= v.size:
            raise ValueError("GpuBVH needs >= 2 triangles whose vertex indices lie in the vertex array")
        self.ntris = idx.size // 3
        self.vertices = DeviceArray.from_host(v)
        self.indices = DeviceArray.from_host(idx)
        self.mesh_ids = None if mesh_ids is None else DeviceArray.from_host(np.ascontiguousarray(mesh_ids, np.int32))
        self.mesh_id = int(mesh_id)
        self.nodes = DeviceArray(STACK_NODE_DTYPE, max(1, self.ntris - 1))
        self.triangles = DeviceArray(TRIANGLE_DTYPE, max(1, self.ntris))
        self.workspace = DeviceArray(np.uint8, max(256, int(lib().crt_build_bvh_gpu_workspace(self.ntris))))
        self.entities = DeviceArray.from_host(_identity_entity(max(0, self.ntris - 1)))
        self.build(stream)
        self.scene = SceneView(FORMAT_STACK, 1, max(0, self.ntris - 1), self.nodes.ptr, self.triangles.ptr,
                               self.vertices.ptr, self.entities.ptr, None)

    def build(self, stream=None) -> None:
        """Rebuilds from the current vertex buffer (asynchronous)."""
        check(lib().crt_build_bvh_gpu(self.vertices.ptr, self.vertices.count, self.indices.ptr, self.ntris,
                                      self.mesh_ids.ptr if self.mesh_ids is not None else None, self.mesh_id,
                                      self.t_offset,
                                      self.nodes.ptr, self.triangles.ptr, self.workspace.ptr, self.workspace.nbytes,
                                      _stream_handle(stream)))


class GpuWide:
    """crt_wide_build_gpu: the wide layout (2-triangle leaves) of a GpuBVH, collapsed on the device and numbered
    breadth-first; `scene` is its crt_scene_view (one identity entity over the wide nodes).

    max_levels=None collapses with crt_wide_build_gpu (reads the node count back; node_count is exact).
    An integer uses crt_wide_build_gpu_async with that many levels: nothing is read back, so build() can be
    captured into a Graph; the scene then spans ntris nodes (the capacity of `groups`, which bounds every
    valid link), and result() reads the node count and status words."""

    def __init__(self, bvh: GpuBVH, width: int = 4, stream=None, max_levels: int | None = None):
        if width not in (4, 8):
            raise ValueError("width must be 4 or 8")
        self.bvh = bvh
        self.width = width
        self.max_levels = max_levels
        self.groups = DeviceArray(WIDE_GROUP_DTYPE, bvh.ntris * (width // 4))
        self.records = DeviceArray(TRI_RECORD_DTYPE, bvh.ntris)
        if max_levels is None:
            ws = lib().crt_wide_build_gpu_workspace(bvh.ntris, width)
        else:
            ws = lib().crt_wide_build_gpu_async_workspace(bvh.ntris, width, int(max_levels))
            self.out = DeviceArray(np.int64, 3)   # [node count, levels needed, status]
        self.workspace = DeviceArray(np.uint8, max(256, int(ws)))
        self.entities = DeviceArray(ENTITY_DTYPE, 1)
        self.node_count = 0
        self.scene = None
        if max_levels is not None:
            self.node_count = bvh.ntris
            self._set_scene()
        self.build(stream)

    def _set_scene(self) -> None:
        b = self.bvh
        self.entities.upload(_identity_entity(self.node_count))
        self.scene = SceneView(FORMAT_WIDE4 if self.width == 4 else FORMAT_WIDE8, 1, self.node_count,
                               self.groups.ptr, b.triangles.ptr, b.vertices.ptr, self.entities.ptr, self.records.ptr)

    def build(self, stream=None) -> None:
        """Collapses the GpuBVH's current tree (synchronizes the stream unless max_levels was given)."""
        b = self.bvh
        if self.max_levels is not None:
            check(lib().crt_wide_build_gpu_async(b.nodes.ptr, b.nodes.count, b.triangles.ptr, b.ntris,
                                                 b.vertices.ptr, b.vertices.count, self.width, int(self.max_levels),
                                                 self.groups.ptr, self.records.ptr, self.out.ptr, self.out.ptr + 16,
                                                 self.workspace.ptr, self.workspace.nbytes, _stream_handle(stream)))
            return
        n = C.c_int64()
        check(lib().crt_wide_build_gpu(b.nodes.ptr, b.nodes.count, b.triangles.ptr, b.ntris, b.vertices.ptr,
                                       b.vertices.count, self.width, self.groups.ptr, self.records.ptr, C.byref(n),
                                       self.workspace.ptr, self.workspace.nbytes, _stream_handle(stream)))
        self.node_count = int(n.value)
        self._set_scene()

    def result(self) -> tuple[int, int]:
        """(wide node count, CRT_WIDE_GPU_* status bits) of the last asynchronous build; synchronizes the device."""
        check(lib().crt_device_sync())
        out = self.out.download()
        self.levels = int(out[1])
        return int(out[0]), int(out[2] & 0xFFFFFFFF)


class DeviceWide:
    """A wide layout uploaded to the device, with its crt_scene_view (scene)."""

    def __init__(self, w: WideBVH, triangles=None, vertices=None):
        """triangles / vertices: the BVH2 scene's device buffers (DeviceArray or device address), needed only
        by crt_hit_attributes."""
        self.host = w
        self.groups = DeviceArray.from_host(w.groups)
        self.records = DeviceArray.from_host(w.records)
        self.entities = DeviceArray.from_host(w.entities)
        addr = lambda b: b.ptr if isinstance(b, DeviceArray) else b
        self.scene = SceneView(w.format, w.entities.size, w.stats["node_count"], self.groups.ptr, addr(triangles),
                               addr(vertices), self.entities.ptr, self.records.ptr)


def upload_wide(src_format: int, host_buffers: dict, width: int = 4, max_leaf: int = 2, triangles=None,
                vertices=None, collapse: int = COLLAPSE_GREEDY, node_cost: float = 0.0,
                tri_cost: float = 0.0) -> DeviceWide:
    """Derives the wide layout from a context's host buffers and uploads it."""
    hb = host_buffers
    return DeviceWide(wide_build(src_format, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], width,
                                 max_leaf, collapse, node_cost, tri_cost), triangles, vertices)


def _colmajor(m) -> np.ndarray:
    """4x4 matrix in math convention (m[row, col]) -> 16 floats column-major (glm layout)."""
    m = np.asarray(m, dtype=np.float32)
    if m.shape == (16,):
        return np.ascontiguousarray(m)
    return np.ascontiguousarray(m.T.reshape(16))


def queue_block() -> DeviceArray:
    """Device memory for one launch's work-queue counters (CRT_QUEUE_BYTES, 256-byte aligned)."""
    return DeviceArray(np.uint8, QUEUE_BYTES)


def trace(scene: SceneView, query: int, rays: DeviceArray, hits: DeviceArray, n: int | None = None, flags: int = 0,
          block_size: int = 0, lds_stack: int = 0, kernel: int = KERNEL_AUTO, variant: int = 0, grid: int = 0,
          tri_tests: DeviceArray | None = None, stream=None, queue: "DeviceArray | None" = None) -> None:
    """crt_trace: closest- or any-hit for every ray of a device ray buffer (asynchronous).  queue: a work-queue
    block (queue_block()) for the persistent kernels instead of one leased from the library's pool (required when
    the launch is captured into a graph)."""
    n = rays.count if n is None else int(n)
    if n > rays.count or n > hits.count or (tri_tests is not None and n > tri_tests.count):
        raise ValueError("ray/hit/counter buffers smaller than n")
    if queue is not None and queue.nbytes < QUEUE_BYTES:
        raise ValueError("queue block smaller than CRT_QUEUE_BYTES")
    cfg = TraceConfig(block_size, lds_stack, kernel, variant, grid, 0, tri_tests.ptr if tri_tests is not None else None,
                      queue.ptr if queue is not None else None)
    check(lib().crt_trace(C.byref(scene), query, flags, rays.ptr, hits.ptr, n, C.byref(cfg), _stream_handle(stream)))


class RayIntersector:
    """RayIntersector<T> (Intersector.h:60-351); fmt selects T (stackless / stack node)."""

    def __init__(self, fmt: int = FORMAT_STACKLESS):
        h = C.c_void_p()
        check(lib().crt_create(fmt, C.byref(h)))
        self.handle = h.value
        self.format = fmt

    def close(self) -> None:
        if self.handle:
            lib().crt_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_object(self, object_id: int, meshes, child_swap: int = SWAP_NONE, seed: int = 0, threads: int = 0,
                   cache_dir: str | None = None) -> int | None:
        """AddObject.  With cache_dir, crt_add_object_cached: returns 1 (loaded), 0 (built and stored) or -1
        (built, not stored)."""
        arr, keep = _meshes(meshes)
        opts = BuildOptions(child_swap, seed & 0xFFFFFFFF, threads)
        if cache_dir is None:
            check(lib().crt_add_object(self.handle, object_id, arr, len(meshes), C.byref(opts)))
            return None
        hit = C.c_int32()
        check(lib().crt_add_object_cached(self.handle, object_id, arr, len(meshes), C.byref(opts),
                                          os.fsencode(cache_dir), C.byref(hit)))
        del keep
        return int(hit.value)

    def add_object_prebuilt(self, object_id: int, nodes: np.ndarray, triangles: np.ndarray, vertices: np.ndarray):
        nodes = np.ascontiguousarray(nodes, dtype=node_dtype(self.format))
        triangles = np.ascontiguousarray(triangles, dtype=TRIANGLE_DTYPE)
        vertices = np.ascontiguousarray(vertices, dtype=VERTEX_DTYPE)
        check(lib().crt_add_object_prebuilt(self.handle, object_id, ptr(nodes), nodes.size, ptr(triangles),
                                            triangles.size, ptr(vertices), vertices.size))

    def push_entity(self, object_id: int, model=None, emissive: float = 0.0, translucency: float = 0.0):
        m = _colmajor(np.eye(4, dtype=np.float32) if model is None else model)
        check(lib().crt_push_entity(self.handle, object_id, ptr(m), float(emissive), float(translucency)))

    def buffer_entities(self):
        check(lib().crt_buffer_entities(self.handle))

    def buffer_data(self, clear_cpu_data: bool = False):
        check(lib().crt_buffer_data(self.handle, int(bool(clear_cpu_data))))

    def host_buffers(self) -> dict:
        p = [C.c_void_p() for _ in range(4)]
        n = [C.c_int64() for _ in range(4)]
        check(lib().crt_context_host_buffers(self.handle, C.byref(p[0]), C.byref(n[0]), C.byref(p[1]),
                                             C.byref(n[1]), C.byref(p[2]), C.byref(n[2]), C.byref(p[3]),
                                             C.byref(n[3])))
        out = {}
        for key, pp, nn, dt in zip(("nodes", "triangles", "vertices", "entities"), p, n,
                                   (node_dtype(self.format), TRIANGLE_DTYPE, VERTEX_DTYPE, ENTITY_DTYPE)):
            cnt = int(nn.value)
            out[key] = (np.frombuffer((C.c_char * (cnt * dt.itemsize)).from_address(pp.value), dtype=dt).copy()
                        if cnt else np.zeros(0, dt))
        return out

    def scene(self) -> SceneView:
        s = SceneView()
        check(lib().crt_context_scene(self.handle, C.byref(s)))
        return s

    def build_wide(self, width: int = 4, max_leaf: int = 2, collapse: int = COLLAPSE_GREEDY, node_cost: float = 0.0,
                   tri_cost: float = 0.0) -> None:
        """crt_context_build_wide(_ex): derive and upload the BVH4 / BVH8 layout of the buffered scene."""
        if collapse == COLLAPSE_GREEDY and node_cost == 0.0 and tri_cost == 0.0:
            check(lib().crt_context_build_wide(self.handle, int(width), int(max_leaf)))
        else:
            opt = WideOptions(int(width), int(max_leaf), int(collapse), float(node_cost), float(tri_cost))
            check(lib().crt_context_build_wide_ex(self.handle, C.byref(opt)))

    def wide_scene(self) -> SceneView:
        s = SceneView()
        check(lib().crt_context_wide_scene(self.handle, C.byref(s)))
        return s

    def intersect(self, query: int, rays: DeviceArray, hits: DeviceArray, flags: int = 0, stream=None):
        check(lib().crt_intersect(self.handle, query, flags, rays.ptr, hits.ptr, rays.count,
                                  _stream_handle(stream)))

    def intersect_primary(self, width: int, height: int, inv_view, inv_proj, flags: int = 0, stream=None):
        rays = DeviceArray(RAY_DTYPE, width * height)
        hits = DeviceArray(HIT_DTYPE, width * height)
        iv, ip = _colmajor(inv_view), _colmajor(inv_proj)
        check(lib().crt_intersect_primary(self.handle, rays.ptr, hits.ptr, width, height, ptr(iv), ptr(ip), flags,
                                          _stream_handle(stream)))
        return rays, hits


class BatchPlan:
    """crt_batch_plan: every (scene, query, rays, hits) batch in one persistent launch.

    batches: list of (SceneView, query, rays DeviceArray, hits DeviceArray[, flags[, tri_tests DeviceArray]])."""

    def __init__(self, batches, config: BatchConfig | None = None):
        self._arr = _batch_array(batches)
        h = C.c_void_p()
        check(lib().crt_batch_plan_create(self._arr, len(batches), C.byref(config) if config else None,
                                          C.byref(h)))
        self.handle = h.value
        self.rays = int(lib().crt_batch_plan_rays(self.handle))

    def launch(self, stream=None, graph: bool = False) -> None:
        fn = lib().crt_batch_plan_launch_graph if graph else lib().crt_batch_plan_launch
        check(fn(self.handle, _stream_handle(stream)))

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().crt_batch_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _batch_array(batches):
    arr = (Batch * max(1, len(batches)))()
    for i, b in enumerate(batches):
        scene, query, rays, hits = b[:4]
        flags = b[4] if len(b) > 4 else 0
        tt = b[5] if len(b) > 5 else None
        arr[i] = Batch(scene, query, flags, rays.ptr, hits.ptr, min(rays.count, hits.count),
                       tt.ptr if tt is not None else None)
    return arr


def autotune_batched(batches, candidates=None, repeats: int = 2, stream=None):
    """crt_autotune_batched: the fastest BatchConfig for these batches and its time per launch (ms)."""
    arr = _batch_array(batches)
    best = BatchConfig()
    ms = C.c_float()
    if candidates:
        carr = (BatchConfig * len(candidates))(*candidates)
        check(lib().crt_autotune_batched(arr, len(batches), carr, len(candidates), repeats, _stream_handle(stream),
                                         C.byref(best), C.byref(ms)))
    else:
        check(lib().crt_autotune_batched(arr, len(batches), None, 0, repeats, _stream_handle(stream), C.byref(best),
                                         C.byref(ms)))
    return best, float(ms.value)


def autotune_batched_layouts(layouts, candidates=None, repeats: int = 2, stream=None):
    """crt_autotune_batched_layouts: `layouts` is a list of batch lists, the same batches over different node
    layouts (BVH2 / BVH4 / BVH8 scene views); returns (index of the fastest layout, its BatchConfig, ms)."""
    n = len(layouts[0])
    if any(len(b) != n for b in layouts):
        raise ValueError("every layout needs the same number of batches")
    arr = _batch_array([b for batches in layouts for b in batches])
    best = BatchConfig()
    ms = C.c_float()
    layout = C.c_int32()
    carr = (BatchConfig * len(candidates))(*candidates) if candidates else None
    check(lib().crt_autotune_batched_layouts(arr, n, len(layouts), carr, len(candidates) if candidates else 0,
                                             repeats, _stream_handle(stream), C.byref(layout), C.byref(best),
                                             C.byref(ms)))
    return int(layout.value), best, float(ms.value)


class ShadingModel:
    """Differentiable Cook-Torrance + one-bounce shading (crt_shade_forward / crt_shade_backward).

    albedo [ntri, 3], roughness [ntri], light_pos [nl, 3], intensity [nl] live on the device."""

    def __init__(self, albedo, roughness, light_pos, intensity):
        self.ntri = int(np.asarray(roughness).size)
        self.nl = int(np.asarray(intensity).size)
        self.d_albedo = DeviceArray.from_host(np.ascontiguousarray(albedo, np.float32).reshape(-1))
        self.d_rough = DeviceArray.from_host(np.ascontiguousarray(roughness, np.float32))
        self.d_lpos = DeviceArray.from_host(np.ascontiguousarray(light_pos, np.float32).reshape(-1))
        self.d_inten = DeviceArray.from_host(np.ascontiguousarray(intensity, np.float32))

    def params(self) -> ShadeParams:
        return ShadeParams(self.d_albedo.ptr, self.d_rough.ptr, self.d_lpos.ptr, self.d_inten.ptr, self.ntri, self.nl)

    def set(self, albedo=None, roughness=None, intensity=None):
        if albedo is not None:
            self.d_albedo.upload(np.ascontiguousarray(albedo, np.float32).reshape(-1))
        if roughness is not None:
            self.d_rough.upload(np.ascontiguousarray(roughness, np.float32))
        if intensity is not None:
            self.d_inten.upload(np.ascontiguousarray(intensity, np.float32))

    def forward(self, scene: SceneView, rays: DeviceArray, hits: DeviceArray, color: DeviceArray,
                records: DeviceArray, seed: int = 4, stream=None) -> None:
        p = self.params()
        check(lib().crt_shade_forward(C.byref(scene), rays.ptr, hits.ptr, rays.count, C.byref(p), seed & 0xFFFFFFFF,
                                      color.ptr, records.ptr, _stream_handle(stream)))

    def backward(self, scene: SceneView, records: DeviceArray, color: DeviceArray, target: DeviceArray,
                 g_albedo: DeviceArray, g_rough: DeviceArray, g_inten: DeviceArray, stream=None) -> None:
        p = self.params()
        check(lib().crt_shade_backward(C.byref(scene), records.ptr, records.count, C.byref(p), color.ptr, target.ptr,
                                       g_albedo.ptr, g_rough.ptr, g_inten.ptr, _stream_handle(stream)))
