"""Python access to the checkers in oracle/ (TEST INFRASTRUCTURE ONLY).

  liboracle.so   C restatement of the reference's GLSL traversal (oracle_trace.c)
  _ref/ref_bvh   the reference's own BVHConstructor.cpp (built from the
                 reference tree by oracle/Makefile when that tree is present)

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; libcandela_rt never does.
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_BVH = os.path.join(HERE, "_ref", "ref_bvh")

VERTEX_DTYPE = np.dtype([("position", "<f4", 4), ("normal_tangent_data", "<u4", 3), ("texcoords", "<u4")])
TRIANGLE_DTYPE = np.dtype([("packed", "<i4", 4)])
STACKLESS_NODE_DTYPE = np.dtype([("min", "<f4", 4), ("max", "<f4", 4)])
BOUNDS_DTYPE = np.dtype([("min", "<f4", 4), ("max", "<f4", 4)])
STACK_NODE_DTYPE = np.dtype([("left", BOUNDS_DTYPE), ("right", BOUNDS_DTYPE)])
ENTITY_DTYPE = np.dtype([("model", "<f4", 16), ("inverse", "<f4", 16), ("node_offset", "<i4"),
                         ("node_count", "<i4"), ("data", "<i4", 14)])
RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmax", "<f4"), ("dir", "<f4", 3), ("id", "<u4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("w", "<f4"), ("tri", "<i4"),
                      ("mesh", "<i4"), ("entity", "<i4"), ("status", "<u4")])
WIDE_GROUP_DTYPE = np.dtype([("lo_x", "<f4", 4), ("hi_x", "<f4", 4), ("lo_y", "<f4", 4), ("hi_y", "<f4", 4),
                             ("lo_z", "<f4", 4), ("hi_z", "<f4", 4), ("child", "<i4", 4), ("count", "<i4", 4)])
TRI_RECORD_DTYPE = np.dtype([("v0", "<f4", 3), ("tri", "<i4"), ("v1", "<f4", 3), ("mesh", "<i4"),
                             ("v2", "<f4", 3), ("pad", "<i4")])
TEXTURE_REF_DTYPE = np.dtype([("model_color", "<f4", 4), ("albedo", "<i4"), ("normal", "<i4"), ("pad", "<i4", 2)])
TEXTURE_DTYPE = np.dtype([("texels", "<u8"), ("width", "<i4"), ("height", "<i4")])
HIT_ATTR_DTYPE = np.dtype([("normal", "<f4", 3), ("emissivity", "<f4"), ("albedo", "<f4", 3), ("alpha", "<f4")])
PATH_PARAMS_DTYPE = np.dtype([("inv_view", "<f4", 16), ("inv_proj", "<f4", 16), ("width", "<i4"), ("height", "<i4"),
                              ("spp", "<i4"), ("seed", "<u4"), ("max_bounces", "<i4"), ("rr", "<f4"), ("tmax", "<f4"),
                              ("sky", "<f4", 3), ("albedo", "<u8"), ("nmesh", "<i4"), ("pad", "<i4")])
PATH_SAMPLE_DTYPE = np.dtype([("radiance", "<f4", 3), ("segments", "<i4"), ("visits", "<u4"), ("tests", "<u4"),
                              ("status", "<u4"), ("pad", "<u4")])

RAYGEN_PARAMS_DTYPE = np.dtype([("seed", "<u4"), ("offset", "<f4"), ("ao_tmax", "<f4"), ("gi_tmax", "<f4"),
                                ("spec_offset", "<f4", 2), ("spec_tmax", "<f4"), ("roughness_scale", "<f4"),
                                ("roughness", "<u8"), ("nmesh", "<i4"), ("pad", "<i4")])
PROBE_GRID_DTYPE = np.dtype([("res", "<i4", 3), ("seed", "<u4"), ("box_origin", "<f4", 3), ("tmax", "<f4"),
                             ("size", "<f4", 3), ("pad", "<i4")])

_lib = None


def available() -> bool:
    return os.path.exists(ORACLE_LIB)


def ref_available() -> bool:
    return os.path.exists(REF_BVH)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not available():
            raise ImportError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        l = C.CDLL(ORACLE_LIB)
        vp, i32, i64, u32, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_float
        l.oracle_trace.restype = C.c_int
        l.oracle_trace.argtypes = [i32, i32, i64, vp, vp, vp, vp, C.c_int, u32, vp, vp, i64, C.c_int, vp, i32]
        l.oracle_trace_paths.restype = None
        l.oracle_trace_paths.argtypes = [vp, i32, i64, vp, vp, vp, i32, vp, vp, i64, vp, i32]
        l.oracle_gen_rays.restype = i64
        l.oracle_gen_rays.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, i64, vp, vp, vp, vp]
        l.oracle_gen_probe_rays.restype = None
        l.oracle_gen_probe_rays.argtypes = [vp, vp]
        l.oracle_build_lbvh.restype = C.c_int
        l.oracle_build_lbvh.argtypes = [vp, vp, i64, vp, i32, i32, vp, vp]
        l.oracle_spread3.restype = u32
        l.oracle_spread3.argtypes = [u32]
        l.oracle_sincos_turn.restype = None
        l.oracle_sincos_turn.argtypes = [f32, vp, vp]
        l.oracle_hit_attributes.restype = None
        l.oracle_hit_attributes.argtypes = [vp, vp, vp, vp, i32, vp, i32, vp, i64, vp]
        l.oracle_half_to_float.restype = f32
        l.oracle_half_to_float.argtypes = [u32]
        l.oracle_sample.restype = None
        l.oracle_sample.argtypes = [vp, i32, i32, f32, f32, vp]
        l.oracle_srgb_table.restype = None
        l.oracle_srgb_table.argtypes = [vp]
        l.oracle_trace_wide.restype = C.c_int
        l.oracle_trace_wide.argtypes = [i32, i32, i64, vp, vp, vp, C.c_int, u32, vp, vp, i64, C.c_int, vp, i32, i32]
        l.oracle_trace_wide_spec.restype = C.c_int
        l.oracle_trace_wide_spec.argtypes = [i32, i32, i64, vp, vp, vp, C.c_int, u32, vp, vp, i64, C.c_int, vp, i32,
                                             i32, i32, i32]
        l.oracle_trace_stackless_spec.restype = C.c_int
        l.oracle_trace_stackless_spec.argtypes = [i32, i64, vp, vp, vp, vp, C.c_int, u32, vp, vp, i64, C.c_int, vp,
                                                  i32, i32, i32]
        l.oracle_trace_stack_spec.restype = C.c_int
        l.oracle_trace_stack_spec.argtypes = [i32, i64, vp, vp, vp, vp, C.c_int, u32, vp, vp, i64, C.c_int, vp, i32,
                                              i32, i32]
        l.oracle_bruteforce_closest.restype = C.c_int
        l.oracle_bruteforce_closest.argtypes = [vp, vp, vp, i64, i64, vp, vp, vp, i64]
        l.oracle_primary_rays.restype = None
        l.oracle_primary_rays.argtypes = [C.c_int, C.c_int, vp, vp, f32, vp]
        l.oracle_barycentrics.restype = None
        l.oracle_barycentrics.argtypes = [vp, vp, vp, vp, vp]
        l.oracle_ray_triangle.restype = None
        l.oracle_ray_triangle.argtypes = [vp, vp, vp, vp, vp, vp]
        l.oracle_bounds_stackless.restype = f32
        l.oracle_bounds_stackless.argtypes = [vp, vp, vp, vp, f32, f32]
        l.oracle_bounds_stack.restype = f32
        l.oracle_bounds_stack.argtypes = [vp, vp, vp, vp, f32]
        f64 = C.c_double
        l.oracle_shade_forward.restype = None
        l.oracle_shade_forward.argtypes = [i64, vp, vp, vp, vp, vp, C.c_int, vp]
        l.oracle_shade_adjoint.restype = f64
        l.oracle_shade_adjoint.argtypes = [i64, vp, vp, vp, vp, vp, C.c_int, vp, C.c_int, vp, vp, vp]
        l.oracle_shade_fd.restype = None
        l.oracle_shade_fd.argtypes = [i64, vp, vp, vp, vp, vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, f64, vp]
        l.oracle_cook_torrance.restype = None
        l.oracle_cook_torrance.argtypes = [vp, vp, vp, vp, f64, f64, vp]
        _lib = l
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def trace(fmt: int, nodes: np.ndarray, tris: np.ndarray, verts: np.ndarray, ents: np.ndarray, query: int,
          rays: np.ndarray, flags: int = 0, threads: int = 0, max_iter: int = 0, with_tests: bool = False):
    """Host traversal of `rays` (RAY_DTYPE) -> hits (HIT_DTYPE) [, triangle tests per ray]."""
    nodes = np.ascontiguousarray(nodes)
    tris = np.ascontiguousarray(tris, TRIANGLE_DTYPE)
    verts = np.ascontiguousarray(verts, VERTEX_DTYPE)
    ents = np.ascontiguousarray(ents, ENTITY_DTYPE)
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.empty(rays.size, HIT_DTYPE)
    tests = np.zeros(rays.size, np.uint32) if with_tests else None
    if threads <= 0:
        threads = min(16, os.cpu_count() or 1)   # at most 16 worker threads
    rc = lib().oracle_trace(fmt, ents.size, nodes.size, _p(nodes), _p(tris), _p(verts), _p(ents), query, flags,
                            _p(rays), _p(hits), rays.size, threads, _p(tests) if tests is not None else None,
                            max_iter)
    if rc != 0:
        raise RuntimeError("oracle_trace failed")
    return (hits, tests) if with_tests else hits


def trace_wide(width: int, groups: np.ndarray, records: np.ndarray, ents: np.ndarray, query: int, rays: np.ndarray,
               flags: int = 0, threads: int = 0, max_iter: int = 0, stack_cap: int = 80, with_tests: bool = False):
    """Host walk of the wide layout (the twin of the GPU wide walk; stack_cap = LDS depth + 64)."""
    groups = np.ascontiguousarray(groups, WIDE_GROUP_DTYPE)
    records = np.ascontiguousarray(records, TRI_RECORD_DTYPE)
    ents = np.ascontiguousarray(ents, ENTITY_DTYPE)
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.empty(rays.size, HIT_DTYPE)
    tests = np.zeros(rays.size, np.uint32) if with_tests else None
    if threads <= 0:
        threads = min(16, os.cpu_count() or 1)
    nodes = groups.size * 4 // width
    rc = lib().oracle_trace_wide(width, ents.size, nodes, _p(groups), _p(records), _p(ents), query, flags, _p(rays),
                                 _p(hits), rays.size, threads, _p(tests) if tests is not None else None, max_iter,
                                 stack_cap)
    if rc != 0:
        raise RuntimeError("oracle_trace_wide failed")
    return (hits, tests) if with_tests else hits


# Schedules of libcandela_rt's speculative wide kernels by variant: (LDS stack entries, inner node steps per round,
# queue slots per lane over the BVH4; the BVH8 kernels use 4 more, at most 16).
SPEC_PRESETS = {0: (8, 4, 8), 1: (8, 8, 8), 2: (8, 2, 8), 3: (8, 4, 8), 4: (8, 8, 8), 5: (7, 6, 6)}
# The refilling speculative kernel's presets as schedules of the same twin.  A lane's sequence of node steps and
# leaf phases does not depend on which rays share its wave, so the grid twin with the refilling preset's
# (LDS depth, inner steps, queue slots) gives that launch's records, counters included.
SPEC_P_PRESETS = {0: (8, 4, 8), 1: (8, 8, 8), 2: (8, 8, 8), 3: (8, 6, 8), 4: (8, 8, 8), 5: (7, 8, 6)}


def spec_queue_slots(width: int, q4: int) -> int:
    return q4 if width == 4 else min(q4 + 4, 16)


def trace_wide_spec(width: int, groups: np.ndarray, records: np.ndarray, ents: np.ndarray, query: int,
                    rays: np.ndarray, variant: int = 0, flags: int = 0, threads: int = 0, max_iter: int = 0,
                    with_tests: bool = False, schedule=None):
    """Host twin of the speculative wide kernel's grid launch (64 consecutive rays per wave), replaying the
    wave schedule of preset `variant` (or `schedule` = (lds_depth, inner, queue slots over the BVH4))."""
    depth, inner, q4 = schedule if schedule is not None else SPEC_PRESETS[variant]
    groups = np.ascontiguousarray(groups, WIDE_GROUP_DTYPE)
    records = np.ascontiguousarray(records, TRI_RECORD_DTYPE)
    ents = np.ascontiguousarray(ents, ENTITY_DTYPE)
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.empty(rays.size, HIT_DTYPE)
    tests = np.zeros(rays.size, np.uint32) if with_tests else None
    if threads <= 0:
        threads = min(16, os.cpu_count() or 1)
    nodes = groups.size * 4 // width
    rc = lib().oracle_trace_wide_spec(width, ents.size, nodes, _p(groups), _p(records), _p(ents), query, flags,
                                      _p(rays), _p(hits), rays.size, threads,
                                      _p(tests) if tests is not None else None, max_iter, depth + 64, inner,
                                      spec_queue_slots(width, q4))
    if rc != 0:
        raise RuntimeError("oracle_trace_wide_spec failed")
    return (hits, tests) if with_tests else hits


# Schedules of libcandela_rt's speculative kernels over the two BVH2 layouts (stackless and stack) by variant:
# (inner node steps per round, queue slots per lane).
SBVH2_PRESETS = {0: (8, 8), 1: (16, 8), 2: (16, 12), 3: (32, 16)}
# ... and of the refilling kernels over those layouts (see SPEC_P_PRESETS)
SBVH2_P_PRESETS = {0: (16, 8), 1: (32, 16), 2: (16, 12), 3: (32, 16)}


def _trace_bvh2_spec(stack, nodes, tris, verts, ents, query, rays, variant, flags, threads, max_iter, with_tests,
                     schedule):
    inner, q = schedule if schedule is not None else SBVH2_PRESETS[variant]
    nodes = np.ascontiguousarray(nodes, STACK_NODE_DTYPE if stack else STACKLESS_NODE_DTYPE)
    tris = np.ascontiguousarray(tris, TRIANGLE_DTYPE)
    verts = np.ascontiguousarray(verts, VERTEX_DTYPE)
    ents = np.ascontiguousarray(ents, ENTITY_DTYPE)
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    hits = np.empty(rays.size, HIT_DTYPE)
    tests = np.zeros(rays.size, np.uint32) if with_tests else None
    if threads <= 0:
        threads = min(16, os.cpu_count() or 1)
    fn = lib().oracle_trace_stack_spec if stack else lib().oracle_trace_stackless_spec
    rc = fn(ents.size, nodes.size, _p(nodes), _p(tris), _p(verts), _p(ents), query, flags, _p(rays), _p(hits),
            rays.size, threads, _p(tests) if tests is not None else None, max_iter, inner, q)
    if rc != 0:
        raise RuntimeError("oracle_trace_%s_spec failed" % ("stack" if stack else "stackless"))
    return (hits, tests) if with_tests else hits


def trace_stackless_spec(nodes: np.ndarray, tris: np.ndarray, verts: np.ndarray, ents: np.ndarray, query: int,
                         rays: np.ndarray, variant: int = 0, flags: int = 0, threads: int = 0, max_iter: int = 0,
                         with_tests: bool = False, schedule=None):
    """Host twin of the speculative stackless kernel's grid launch (64 consecutive rays per wave), replaying the
    wave schedule of preset `variant` (or `schedule` = (inner, queue slots))."""
    return _trace_bvh2_spec(False, nodes, tris, verts, ents, query, rays, variant, flags, threads, max_iter,
                            with_tests, schedule)


def trace_stack_spec(nodes: np.ndarray, tris: np.ndarray, verts: np.ndarray, ents: np.ndarray, query: int,
                     rays: np.ndarray, variant: int = 0, flags: int = 0, threads: int = 0, max_iter: int = 0,
                     with_tests: bool = False, schedule=None):
    """Host twin of the speculative stack-layout kernel's grid launch, as trace_stackless_spec (queue slots
    2..16)."""
    return _trace_bvh2_spec(True, nodes, tris, verts, ents, query, rays, variant, flags, threads, max_iter,
                            with_tests, schedule)


def trace_paths(params: np.ndarray, width: int, groups, records, ents, albedo, paths, stack_cap: int = 72):
    """The path kernel's per-path sequence on the host, for the path indices `paths`.
    params: one PATH_PARAMS_DTYPE record (its albedo pointer is ignored; `albedo` [nmesh, 3] is used)."""
    params = np.ascontiguousarray(params, PATH_PARAMS_DTYPE).reshape(1).copy()
    albedo = np.ascontiguousarray(albedo, np.float32)
    params["nmesh"] = albedo.size // 3
    groups = np.ascontiguousarray(groups, WIDE_GROUP_DTYPE)
    records = np.ascontiguousarray(records, TRI_RECORD_DTYPE)
    ents = np.ascontiguousarray(ents, ENTITY_DTYPE)
    paths = np.ascontiguousarray(paths, np.int64)
    out = np.empty(paths.size, PATH_SAMPLE_DTYPE)
    lib().oracle_trace_paths(_p(params), width, groups.size * 4 // width, _p(groups), _p(records), _p(ents), ents.size,
                             _p(albedo), _p(paths), paths.size, _p(out), stack_cap)
    return out


def sincos_turn(u: float):
    s, c = C.c_float(), C.c_float()
    lib().oracle_sincos_turn(float(u), C.byref(s), C.byref(c))
    return s.value, c.value


def hit_attributes(tris, verts, ents, refs, textures, hits):
    """GetData on the host.  textures: list of (texels uint32 [h, w] or [h*w], width, height)."""
    tris = np.ascontiguousarray(tris, TRIANGLE_DTYPE)
    verts = np.ascontiguousarray(verts, VERTEX_DTYPE)
    ents = np.ascontiguousarray(ents, ENTITY_DTYPE)
    refs = np.ascontiguousarray(refs, TEXTURE_REF_DTYPE)
    hits = np.ascontiguousarray(hits, HIT_DTYPE)
    keep = [np.ascontiguousarray(t[0], np.uint32) for t in textures]
    desc = np.zeros(max(1, len(textures)), TEXTURE_DTYPE)
    for k, (t, (_, w, h)) in enumerate(zip(keep, textures)):
        desc[k] = (t.ctypes.data, w, h)
    out = np.empty(hits.size, HIT_ATTR_DTYPE)
    lib().oracle_hit_attributes(_p(tris), _p(verts), _p(ents), _p(refs), refs.size, _p(desc), len(textures),
                                _p(hits), hits.size, _p(out))
    return out


def half_to_float(h: int) -> float:
    return float(lib().oracle_half_to_float(int(h)))


def sample(texels, width: int, height: int, u: float, v: float) -> np.ndarray:
    t = np.ascontiguousarray(texels, np.uint32)
    out = np.empty(3, np.float32)
    lib().oracle_sample(_p(t), width, height, float(u), float(v), _p(out))
    return out


def srgb_table() -> np.ndarray:
    out = np.empty(256, np.float32)
    lib().oracle_srgb_table(_p(out))
    return out


def bruteforce_closest(tris, verts, ent, rays, tri_begin: int = 0, tri_end: int | None = None):
    tris = np.ascontiguousarray(tris, TRIANGLE_DTYPE)
    verts = np.ascontiguousarray(verts, VERTEX_DTYPE)
    ent = np.ascontiguousarray(np.atleast_1d(ent), ENTITY_DTYPE)
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    tri_end = tris.size if tri_end is None else tri_end
    t = np.empty(rays.size, np.float32)
    tri = np.empty(rays.size, np.int32)
    lib().oracle_bruteforce_closest(_p(tris), _p(verts), _p(ent), tri_begin, tri_end, _p(rays), _p(t), _p(tri),
                                    rays.size)
    return t, tri


def primary_rays(width: int, height: int, inv_view_colmajor: np.ndarray, inv_proj_colmajor: np.ndarray,
                 tmax: float = 1e6) -> np.ndarray:
    iv = np.ascontiguousarray(inv_view_colmajor, np.float32)
    ip = np.ascontiguousarray(inv_proj_colmajor, np.float32)
    rays = np.empty(width * height, RAY_DTYPE)
    lib().oracle_primary_rays(width, height, _p(iv), _p(ip), tmax, _p(rays))
    return rays


def ray_triangle(ro, rd, a, b, c):
    arrs = [np.ascontiguousarray(x, np.float32) for x in (ro, rd, a, b, c)]
    out = np.empty(3, np.float32)
    lib().oracle_ray_triangle(*[_p(x) for x in arrs], _p(out))
    return out


def barycentrics(p, a, b, c) -> np.ndarray:
    """ComputeBarycentrics (Include/TraverseBVHStackless.glsl:156-172): (u, v, w)."""
    arrs = [np.ascontiguousarray(x, np.float32) for x in (p, a, b, c)]
    out = np.empty(3, np.float32)
    lib().oracle_barycentrics(*[_p(x) for x in arrs], _p(out))
    return out


def bounds_stackless(bmin, bmax, ro, inv, t_min, t_max) -> float:
    arrs = [np.ascontiguousarray(x, np.float32) for x in (bmin, bmax, ro, inv)]
    return float(np.float32(lib().oracle_bounds_stackless(*[_p(x) for x in arrs], t_min, t_max)))


def bounds_stack(bmin, bmax, ro, inv, maxt) -> float:
    arrs = [np.ascontiguousarray(x, np.float32) for x in (bmin, bmax, ro, inv)]
    return float(np.float32(lib().oracle_bounds_stack(*[_p(x) for x in arrs], maxt)))


def identity_entity(node_count: int, node_offset: int = 0) -> np.ndarray:
    e = np.zeros(1, ENTITY_DTYPE)
    eye = np.eye(4, dtype=np.float32).T.reshape(16)
    e["model"] = eye
    e["inverse"] = eye
    e["node_offset"] = node_offset
    e["node_count"] = node_count
    e["data"][0, 1] = np.float32(1.0).view(np.int32)
    return e


def ref_build(fmt: int, verts: np.ndarray, indices: np.ndarray, mesh_ids: np.ndarray):
    """Runs the reference constructor (one process per build) -> (nodes, triangles, info)."""
    if not ref_available():
        raise FileNotFoundError(f"{REF_BVH} not built (needs the reference tree)")
    verts = np.ascontiguousarray(verts, VERTEX_DTYPE)
    indices = np.ascontiguousarray(indices, np.uint32).ravel()
    mesh_ids = np.ascontiguousarray(mesh_ids, np.int32).ravel()
    assert mesh_ids.size == indices.size // 3
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<II", verts.size, indices.size))
            f.write(verts.tobytes())
            f.write(indices.tobytes())
            f.write(mesh_ids.tobytes())
        mode = "stackless" if fmt == 0 else "stack"
        subprocess.run([REF_BVH, mode, fin, fout], check=True, stdout=subprocess.DEVNULL)
        raw = open(fout, "rb").read()
    hdr = struct.unpack("<6Q", raw[:48])
    n_nodes, n_tri = hdr[0], hdr[1]
    nd = STACKLESS_NODE_DTYPE if fmt == 0 else STACK_NODE_DTYPE
    off = 48
    nodes = np.frombuffer(raw, nd, n_nodes, off).copy()
    off += n_nodes * nd.itemsize
    tris = np.frombuffer(raw, TRIANGLE_DTYPE, n_tri, off).copy()
    return nodes, tris, {"array_length": hdr[0], "triangle_count": hdr[1], "node_count": hdr[2],
                         "leaf_count": hdr[3]}


SHADE_RECORD_DTYPE = np.dtype([("p", "<f4", 3), ("tri0", "<i4"), ("n", "<f4", 3), ("tri1", "<i4"),
                               ("eye", "<f4", 3), ("vis", "<u4"), ("q", "<f4", 3), ("pad0", "<f4"),
                               ("nq", "<f4", 3), ("pad1", "<f4")])


def _d(a):
    return np.ascontiguousarray(a, np.float64)


def shade_forward(rec, albedo, rough, lpos, inten):
    """FP64 colors [n, 3] of the records (geometry and visibility from the records)."""
    rec = np.ascontiguousarray(rec, SHADE_RECORD_DTYPE)
    a, r, lp, it = _d(albedo), _d(rough), _d(lpos), _d(inten)
    out = np.empty((rec.size, 3), np.float64)
    lib().oracle_shade_forward(rec.size, _p(rec), _p(a), _p(r), _p(lp), _p(it), it.size, _p(out))
    return out


def shade_adjoint(rec, albedo, rough, lpos, inten, target):
    """(loss, d/d albedo [ntri*3], d/d roughness [ntri], d/d intensity [nl]) in FP64."""
    rec = np.ascontiguousarray(rec, SHADE_RECORD_DTYPE)
    a, r, lp, it = _d(albedo), _d(rough), _d(lpos), _d(inten)
    t = np.ascontiguousarray(target, np.float32)
    ntri = r.size
    ga, gr, gi = np.empty(3 * ntri), np.empty(ntri), np.empty(it.size)
    loss = lib().oracle_shade_adjoint(rec.size, _p(rec), _p(a), _p(r), _p(lp), _p(it), it.size, _p(t), ntri,
                                      _p(ga), _p(gr), _p(gi))
    return float(loss), ga, gr, gi


def shade_fd(rec, albedo, rough, lpos, inten, target, kind, index, h=1e-3):
    """Central finite differences of the FP64 loss for parameters (kind, index)."""
    rec = np.ascontiguousarray(rec, SHADE_RECORD_DTYPE)
    a, r, lp, it = _d(albedo), _d(rough), _d(lpos), _d(inten)
    t = np.ascontiguousarray(target, np.float32)
    kind = np.ascontiguousarray(kind, np.int32)
    index = np.ascontiguousarray(index, np.int32)
    out = np.empty(kind.size, np.float64)
    lib().oracle_shade_fd(rec.size, _p(rec), _p(a), _p(r), _p(lp), _p(it), it.size, _p(t), r.size, _p(kind),
                          _p(index), kind.size, h, _p(out))
    return out


def cook_torrance(x, n, eye, light, intensity, r):
    """(A, kd, spec, dspec/dr) of one light, FP64."""
    arrs = [_d(v) for v in (x, n, eye, light)]
    out = np.empty(4, np.float64)
    lib().oracle_cook_torrance(*[_p(v) for v in arrs], float(intensity), float(r), _p(out))
    return out


def _gen(kind, params, primary, hits, ents, records=None, tris=None, verts=None, roughness=None):
    params = np.ascontiguousarray(params, RAYGEN_PARAMS_DTYPE).reshape(1).copy()
    primary = np.ascontiguousarray(primary, RAY_DTYPE)
    hits = np.ascontiguousarray(hits, HIT_DTYPE)
    ents = np.ascontiguousarray(ents, ENTITY_DTYPE)
    rough = None if roughness is None else np.ascontiguousarray(roughness, np.float32)
    params["nmesh"] = 0 if rough is None else rough.size
    recs = None if records is None else np.ascontiguousarray(records, TRI_RECORD_DTYPE)
    tr = None if tris is None else np.ascontiguousarray(tris, TRIANGLE_DTYPE)
    vt = None if verts is None else np.ascontiguousarray(verts, VERTEX_DTYPE)
    out0 = np.zeros(hits.size, RAY_DTYPE)
    out1 = np.zeros(hits.size, RAY_DTYPE)
    k = lib().oracle_gen_rays(kind, _p(recs) if recs is not None else None, _p(tr) if tr is not None else None,
                              _p(vt) if vt is not None else None, _p(ents), _p(primary), _p(hits), hits.size,
                              _p(params), _p(rough) if rough is not None else None, _p(out0), _p(out1))
    return out0[:k], out1[:k]


def gen_diffuse_rays(params, primary, hits, ents, records=None, tris=None, verts=None):
    """(AO rays, GI rays) of the records with t > 0, in input order (raygen.hip's host twin)."""
    return _gen(0, params, primary, hits, ents, records, tris, verts)


def gen_specular_rays(params, primary, hits, ents, records=None, tris=None, verts=None, roughness=None):
    """Reflection rays of the records with t > 0, in input order."""
    return _gen(1, params, primary, hits, ents, records, tris, verts, roughness)[0]


def gen_probe_rays(grid) -> np.ndarray:
    grid = np.ascontiguousarray(grid, PROBE_GRID_DTYPE).reshape(1)
    out = np.zeros(int(np.prod(grid["res"][0].astype(np.int64))), RAY_DTYPE)
    lib().oracle_gen_probe_rays(_p(grid), _p(out))
    return out


def build_lbvh(vertices, indices, mesh_ids=None, mesh_id: int = 0, t_offset: int = 0):
    """The device builder's host twin: (stack nodes [n - 1], triangles [n]) for n = len(indices) // 3."""
    v = np.ascontiguousarray(vertices, VERTEX_DTYPE)
    idx = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    n = idx.size // 3
    m = None if mesh_ids is None else np.ascontiguousarray(mesh_ids, np.int32)
    nodes = np.zeros(max(0, n - 1), STACK_NODE_DTYPE)
    tris = np.zeros(n, TRIANGLE_DTYPE)
    rc = lib().oracle_build_lbvh(_p(v), _p(idx), n, _p(m) if m is not None else None, mesh_id, t_offset, _p(nodes),
                                 _p(tris))
    if rc != 0:
        raise ValueError("oracle_build_lbvh failed (needs >= 2 triangles)")
    return nodes, tris


def spread3(v: int) -> int:
    return int(lib().oracle_spread3(v))
