"""ctypes binding of libcandela_rt.so (the C ABI declared in include/candela_rt.h).

The library is the product: there is no Python or CPU fallback.  Loading fails
loudly when the shared object has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcandela_rt.so")

# --- constants (include/candela_rt.h) -----------------------------------------
CRT_OK = 0
FORMAT_STACKLESS = 0
FORMAT_STACK = 1
FORMAT_WIDE4 = 2
FORMAT_WIDE8 = 3
QUERY_CLOSEST = 0
QUERY_ANY = 1
FLAG_IGNORE_TRANSPARENT = 1 << 0
FLAG_TRI0_IS_MISS = 1 << 1
FLAG_SLOT_ORDER = 1 << 2
STATUS_ITER_CAP = 1 << 24
STATUS_STACK_OVERFLOW = 1 << 25
STATUS_BAD_INDEX = 1 << 26
STATUS_VISITS_MASK = 0x00FFFFFF
SWAP_NONE = 0
SWAP_SEEDED = 1
E_INVALID, E_HIP, E_NOMEM, E_STATE, E_NODEVICE, E_IO, E_CORRUPT, E_STALE = -1, -2, -3, -4, -5, -6, -7, -8
KERNEL_AUTO = 0
KERNEL_THREAD_PER_RAY = 1
KERNEL_PERSISTENT = 2
KERNEL_WAVE = 3
KERNEL_SPECULATIVE, KERNEL_SPECULATIVE_REFILL = 4, 5
WIDE_GPU_BAD_LINK, WIDE_GPU_NOT_TREE, WIDE_GPU_BAD_VERTEX, WIDE_GPU_TOO_DEEP, WIDE_GPU_BAD_LEAF = 1, 2, 4, 8, 16
WIDE_GPU_MAX_LEVELS = 4096

# --- record dtypes (byte layouts of the reference buffers) ---------------------
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
assert WIDE_GROUP_DTYPE.itemsize == 128 and TRI_RECORD_DTYPE.itemsize == 48
TEXTURE_REF_DTYPE = np.dtype([("model_color", "<f4", 4), ("albedo", "<i4"), ("normal", "<i4"), ("pad", "<i4", 2)])
TEXTURE_DTYPE = np.dtype([("texels", "<u8"), ("width", "<i4"), ("height", "<i4")])
HIT_ATTR_DTYPE = np.dtype([("normal", "<f4", 3), ("emissivity", "<f4"), ("albedo", "<f4", 3), ("alpha", "<f4")])
assert TEXTURE_REF_DTYPE.itemsize == 32 and TEXTURE_DTYPE.itemsize == 16 and HIT_ATTR_DTYPE.itemsize == 32
PATH_PARAMS_DTYPE = np.dtype([("inv_view", "<f4", 16), ("inv_proj", "<f4", 16), ("width", "<i4"), ("height", "<i4"),
                              ("spp", "<i4"), ("seed", "<u4"), ("max_bounces", "<i4"), ("rr", "<f4"), ("tmax", "<f4"),
                              ("sky", "<f4", 3), ("albedo", "<u8"), ("nmesh", "<i4"), ("pad", "<i4")])
PATH_SAMPLE_DTYPE = np.dtype([("radiance", "<f4", 3), ("segments", "<i4"), ("visits", "<u4"), ("tests", "<u4"),
                              ("status", "<u4"), ("pad", "<u4")])
assert PATH_PARAMS_DTYPE.itemsize == 184 and PATH_SAMPLE_DTYPE.itemsize == 32
assert VERTEX_DTYPE.itemsize == 32 and TRIANGLE_DTYPE.itemsize == 16
assert STACKLESS_NODE_DTYPE.itemsize == 32 and STACK_NODE_DTYPE.itemsize == 64
assert ENTITY_DTYPE.itemsize == 192 and RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 32
RAYGEN_PARAMS_DTYPE = np.dtype([("seed", "<u4"), ("offset", "<f4"), ("ao_tmax", "<f4"), ("gi_tmax", "<f4"),
                                ("spec_offset", "<f4", 2), ("spec_tmax", "<f4"), ("roughness_scale", "<f4"),
                                ("roughness", "<u8"), ("nmesh", "<i4"), ("pad", "<i4")])
PROBE_GRID_DTYPE = np.dtype([("res", "<i4", 3), ("seed", "<u4"), ("box_origin", "<f4", 3), ("tmax", "<f4"),
                             ("size", "<f4", 3), ("pad", "<i4")])


def node_dtype(fmt: int) -> np.dtype:
    return STACKLESS_NODE_DTYPE if fmt == FORMAT_STACKLESS else STACK_NODE_DTYPE


# --- ctypes structs ---------------------------------------------------------------
class Mesh(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("vertex_count", C.c_uint32), ("indices", C.c_void_p),
                ("index_count", C.c_uint32), ("global_mesh_number", C.c_int32)]


class BuildOptions(C.Structure):
    _fields_ = [("child_swap", C.c_int32), ("seed", C.c_uint32), ("threads", C.c_int32)]


class BvhStats(C.Structure):
    _fields_ = [("triangle_count", C.c_uint64), ("node_count", C.c_uint64), ("leaf_count", C.c_uint64),
                ("split_fails", C.c_uint64), ("max_depth", C.c_uint32), ("array_length", C.c_uint64),
                ("vertex_count", C.c_uint64)]


class CacheHeader(C.Structure):
    _fields_ = [("magic", C.c_char * 8), ("header_bytes", C.c_uint32), ("version", C.c_uint32),
                ("format", C.c_int32), ("t_offset", C.c_int32), ("key", C.c_uint64),
                ("triangle_count", C.c_uint64), ("array_length", C.c_uint64), ("vertex_count", C.c_uint64),
                ("node_count", C.c_uint64), ("leaf_count", C.c_uint64), ("split_fails", C.c_uint64),
                ("max_depth", C.c_uint32), ("child_swap", C.c_int32), ("seed", C.c_uint32), ("bins", C.c_uint32),
                ("max_leaf", C.c_uint32), ("reserved0", C.c_uint32), ("section_offset", C.c_uint64 * 3),
                ("section_bytes", C.c_uint64 * 3), ("section_hash", C.c_uint64 * 3), ("toolchain", C.c_char * 32),
                ("reserved1", C.c_uint8 * 40), ("header_hash", C.c_uint64)]


class SceneView(C.Structure):
    _fields_ = [("format", C.c_int32), ("entity_count", C.c_int32), ("total_nodes", C.c_int64),
                ("nodes", C.c_void_p), ("triangles", C.c_void_p), ("vertices", C.c_void_p),
                ("entities", C.c_void_p), ("tri_records", C.c_void_p)]


class WideStats(C.Structure):
    _fields_ = [("node_count", C.c_int64), ("leaf_count", C.c_int64), ("record_count", C.c_int64),
                ("width", C.c_int32), ("max_depth", C.c_int32), ("mean_fill", C.c_double)]


class WideOptions(C.Structure):
    _fields_ = [("width", C.c_int32), ("max_leaf", C.c_int32), ("collapse", C.c_int32), ("node_cost", C.c_float),
                ("tri_cost", C.c_float)]


COLLAPSE_GREEDY, COLLAPSE_SAH = 0, 1


class TraceConfig(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("lds_stack", C.c_int32), ("kernel", C.c_int32), ("variant", C.c_int32),
                ("grid", C.c_int32), ("reserved", C.c_int32), ("tri_tests", C.c_void_p), ("queue", C.c_void_p)]


QUEUE_BYTES = 2048   # CRT_QUEUE_BYTES


class Batch(C.Structure):
    _fields_ = [("scene", SceneView), ("query", C.c_int32), ("flags", C.c_uint32), ("rays", C.c_void_p),
                ("hits", C.c_void_p), ("count", C.c_int64), ("tri_tests", C.c_void_p)]


class BatchConfig(C.Structure):
    _fields_ = [("variant", C.c_int32), ("tiles_per_claim", C.c_int32), ("waves_per_group", C.c_int32),
                ("grid", C.c_int32), ("order", C.c_int32), ("shards", C.c_int32), ("child_order", C.c_int32)]


class ShadeParams(C.Structure):
    _fields_ = [("albedo", C.c_void_p), ("roughness", C.c_void_p), ("light_pos", C.c_void_p),
                ("intensity", C.c_void_p), ("ntri", C.c_int32), ("nlights", C.c_int32)]


SHADE_RECORD_DTYPE = np.dtype([("p", "<f4", 3), ("tri0", "<i4"), ("n", "<f4", 3), ("tri1", "<i4"),
                               ("eye", "<f4", 3), ("vis", "<u4"), ("q", "<f4", 3), ("pad0", "<f4"),
                               ("nq", "<f4", 3), ("pad1", "<f4")])
assert SHADE_RECORD_DTYPE.itemsize == 80


class CrtError(RuntimeError):
    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


_lib = None

# name -> (restype, argtypes); the list is also the export check of the tests.
SIGNATURES = {
    "crt_last_error": (C.c_char_p, []),
    "crt_abi_version": (C.c_int, []),
    "crt_device_count": (C.c_int, []),
    "crt_build_bvh": (C.c_int, [C.c_int, C.c_void_p, C.c_int, C.c_int32, C.c_void_p, C.POINTER(C.c_void_p)]),
    "crt_bvh_free": (None, [C.c_void_p]),
    "crt_bvh_get_stats": (C.c_int, [C.c_void_p, C.POINTER(BvhStats)]),
    "crt_bvh_nodes": (C.c_void_p, [C.c_void_p]),
    "crt_bvh_triangles": (C.c_void_p, [C.c_void_p]),
    "crt_bvh_vertices": (C.c_void_p, [C.c_void_p]),
    "crt_bvh_format": (C.c_int, [C.c_void_p]),
    "crt_bvh_t_offset": (C.c_int32, [C.c_void_p]),
    "crt_raygen_workspace_bytes": (C.c_size_t, [C.c_int64]),
    "crt_build_bvh_gpu_workspace": (C.c_size_t, [C.c_int64]),
    "crt_wide_build_gpu_workspace": (C.c_size_t, [C.c_int64, C.c_int32]),
    "crt_wide_build_gpu": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int32,
                                     C.c_void_p, C.c_void_p, C.POINTER(C.c_int64), C.c_void_p, C.c_size_t,
                                     C.c_void_p]),
    "crt_wide_build_gpu_async_workspace": (C.c_size_t, [C.c_int64, C.c_int32, C.c_int32]),
    "crt_wide_build_gpu_async": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                           C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_size_t, C.c_void_p]),
    "crt_build_bvh_gpu": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int32, C.c_int32,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "crt_gen_diffuse_rays": (C.c_int, [C.POINTER(SceneView), C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "crt_make_entity": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_int32, C.c_int32, C.c_void_p]),
    "crt_gen_specular_rays": (C.c_int, [C.POINTER(SceneView), C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "crt_gen_probe_rays": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "crt_hash64": (C.c_uint64, [C.c_void_p, C.c_size_t, C.c_uint64]),
    "crt_bvh_key": (C.c_int, [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_uint64)]),
    "crt_bvh_save": (C.c_int, [C.c_void_p, C.c_uint64, C.c_char_p]),
    "crt_bvh_load": (C.c_int, [C.c_char_p, C.c_uint64, C.POINTER(C.c_void_p), C.POINTER(CacheHeader)]),
    "crt_bvh_rebase": (C.c_int, [C.c_void_p, C.c_int32]),
    "crt_add_object_cached": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int, C.c_void_p, C.c_char_p,
                                        C.POINTER(C.c_int32)]),
    "crt_device_alloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "crt_device_free": (C.c_int, [C.c_void_p]),
    "crt_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "crt_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "crt_memset_d": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]),
    "crt_device_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "crt_stream_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "crt_stream_destroy": (C.c_int, [C.c_void_p]),
    "crt_stream_sync": (C.c_int, [C.c_void_p]),
    "crt_device_sync": (C.c_int, []),
    "crt_graph_begin": (C.c_int, [C.c_void_p]),
    "crt_graph_end": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "crt_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "crt_graph_destroy": (C.c_int, [C.c_void_p]),
    "crt_event_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "crt_event_destroy": (C.c_int, [C.c_void_p]),
    "crt_event_record": (C.c_int, [C.c_void_p, C.c_void_p]),
    "crt_stream_wait_event": (C.c_int, [C.c_void_p, C.c_void_p]),
    "crt_event_elapsed_ms": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    "crt_trace": (C.c_int, [C.POINTER(SceneView), C.c_int, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int64,
                            C.POINTER(TraceConfig), C.c_void_p]),
    "crt_trace_noalloc": (C.c_int, [C.POINTER(SceneView), C.c_int, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int64,
                                    C.POINTER(TraceConfig), C.c_void_p]),
    "crt_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "crt_destroy": (None, [C.c_void_p]),
    "crt_context_format": (C.c_int, [C.c_void_p]),
    "crt_add_object": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int, C.c_void_p]),
    "crt_add_object_prebuilt": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                          C.c_void_p, C.c_int64]),
    "crt_push_entity": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_float, C.c_float]),
    "crt_buffer_entities": (C.c_int, [C.c_void_p]),
    "crt_buffer_data": (C.c_int, [C.c_void_p, C.c_int]),
    "crt_context_host_buffers": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_int64),
                                           C.POINTER(C.c_void_p), C.POINTER(C.c_int64), C.POINTER(C.c_void_p),
                                           C.POINTER(C.c_int64), C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]),
    "crt_context_scene": (C.c_int, [C.c_void_p, C.POINTER(SceneView)]),
    "crt_intersect": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "crt_intersect_primary": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                        C.c_void_p, C.c_uint32, C.c_void_p]),
    "crt_batch_plan_create": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(BatchConfig), C.POINTER(C.c_void_p)]),
    "crt_batch_plan_destroy": (None, [C.c_void_p]),
    "crt_batch_plan_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "crt_batch_plan_launch_graph": (C.c_int, [C.c_void_p, C.c_void_p]),
    "crt_batch_plan_rays": (C.c_int64, [C.c_void_p]),
    "crt_shade_forward": (C.c_int, [C.POINTER(SceneView), C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(ShadeParams),
                                    C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "crt_shade_backward": (C.c_int, [C.POINTER(SceneView), C.c_void_p, C.c_int64, C.POINTER(ShadeParams), C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "crt_autotune_batched": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                       C.POINTER(BatchConfig), C.POINTER(C.c_float)]),
    "crt_autotune_batched_layouts": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                               C.c_void_p, C.POINTER(C.c_int32), C.POINTER(BatchConfig),
                                               C.POINTER(C.c_float)]),
    "crt_wide_build": (C.c_int, [C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                 C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "crt_wide_build_ex": (C.c_int, [C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                    C.c_void_p, C.c_int32, C.POINTER(WideOptions), C.POINTER(C.c_void_p)]),
    "crt_wide_free": (None, [C.c_void_p]),
    "crt_wide_get_stats": (C.c_int, [C.c_void_p, C.POINTER(WideStats)]),
    "crt_wide_groups": (C.c_void_p, [C.c_void_p]),
    "crt_wide_records": (C.c_void_p, [C.c_void_p]),
    "crt_wide_entities": (C.c_void_p, [C.c_void_p]),
    "crt_context_build_wide": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "crt_context_build_wide_ex": (C.c_int, [C.c_void_p, C.POINTER(WideOptions)]),
    "crt_hit_attributes": (C.c_int, [C.POINTER(SceneView), C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p,
                                     C.c_int64, C.c_void_p, C.c_void_p]),
    "crt_trace_paths": (C.c_int, [C.POINTER(SceneView), C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(TraceConfig),
                                  C.c_void_p]),
    "crt_context_wide_scene": (C.c_int, [C.c_void_p, C.POINTER(SceneView)]),
}


def lib() -> C.CDLL:
    """Loads libcandela_rt.so once; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m candela_amd.build` "
                              "(there is no fallback implementation)")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(rc: int) -> None:
    if rc != CRT_OK:
        msg = lib().crt_last_error()
        raise CrtError(f"libcandela_rt error {rc}: {msg.decode() if msg else ''}", rc)


def ptr(a: np.ndarray) -> int:
    return a.ctypes.data
