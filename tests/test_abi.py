"""The C ABI library loads, exports what include/candela_rt.h declares and rejects bad input (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def header_functions():
    text = open(os.path.join(ROOT, "include", "candela_rt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(crt_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported(crt):
    import candela_amd._lib as L
    names = header_functions()
    assert len(names) >= 30
    lib = crt.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # Every declared function has a ctypes signature in the binding.
    assert sorted(L.SIGNATURES) == names


def test_abi_version(crt):
    assert crt.lib().crt_abi_version() == 2


def test_struct_sizes(crt):
    assert crt.VERTEX_DTYPE.itemsize == 32
    assert crt.TRIANGLE_DTYPE.itemsize == 16
    assert crt.STACKLESS_NODE_DTYPE.itemsize == 32
    assert crt.STACK_NODE_DTYPE.itemsize == 64
    assert crt.ENTITY_DTYPE.itemsize == 192
    assert crt.RAY_DTYPE.itemsize == 32
    assert crt.HIT_DTYPE.itemsize == 32
    import ctypes
    assert ctypes.sizeof(crt.CacheHeader) == 256 and crt.CacheHeader.header_hash.offset == 248
    assert ctypes.sizeof(crt.BatchConfig) == 28 and crt.BatchConfig.child_order.offset == 24
    assert (crt.FLAG_IGNORE_TRANSPARENT, crt.FLAG_TRI0_IS_MISS, crt.FLAG_SLOT_ORDER) == (1, 2, 4)


def test_errors_without_gpu(crt):
    lib = crt.lib()
    h = C.c_void_p()
    assert lib.crt_build_bvh(7, None, 0, 0, None, C.byref(h)) == -1
    assert b"format" in lib.crt_last_error()
    assert lib.crt_create(9, C.byref(h)) == -1
    # trace with a null scene is rejected before any device work
    assert lib.crt_trace(None, 0, 0, None, None, 0, None, None) == -1
    s = crt.SceneView(0, 1, 10, None, None, None, None)
    assert lib.crt_trace(C.byref(s), 0, 0, None, None, 1, None, None) == -1
    s2 = crt.SceneView(0, 0, 0, None, None, None, None)
    cfg = crt.TraceConfig(100, 0)
    assert lib.crt_trace(C.byref(s2), 0, 0, None, None, 0, C.byref(cfg), None) == -1
    assert b"block_size" in lib.crt_last_error()


def test_context_host_side(crt):
    from candela_amd import scenes
    r = crt.RayIntersector(crt.FORMAT_STACKLESS)
    with pytest.raises(crt.CrtError):
        r.push_entity(5)          # object 5 was never added (Intersector.h:203-205)
    r.add_object(0, scenes.boxes_scene())
    r.add_object(1, scenes.boxes_scene(seed=7))
    r.push_entity(0)
    r.push_entity(1, np.diag([2.0, 2.0, 2.0, 1.0]))
    hb = r.host_buffers()
    b0 = crt.build_bvh(crt.FORMAT_STACKLESS, scenes.boxes_scene())
    b1 = crt.build_bvh(crt.FORMAT_STACKLESS, scenes.boxes_scene(seed=7), t_offset=b0.triangles.size)
    n0 = b0.nodes.size
    assert hb["nodes"].size == n0 + b1.nodes.size
    assert hb["nodes"][:n0].tobytes() == b0.nodes.tobytes()
    assert hb["nodes"][n0:].tobytes() == b1.nodes.tobytes()   # node links stay object-local
    # vertex indices of object 1 are offset by object 0's vertex count (Intersector.h:190-197)
    t1 = hb["triangles"][b0.triangles.size:]["packed"]
    assert np.array_equal(t1[:, :3], b1.triangles["packed"][:, :3] + b0.vertices.size)
    assert np.array_equal(t1[:, 3], b1.triangles["packed"][:, 3])
    with pytest.raises(crt.CrtError):
        r.scene()   # nothing buffered yet


def test_device_build_workspaces_and_arguments(crt):
    """Workspace sizes and argument checks of the device builder and both device collapses: host-side only,
    rejected before any device work."""
    lib = crt.lib()
    n = 100000
    assert lib.crt_build_bvh_gpu_workspace(1) == 0
    assert lib.crt_build_bvh_gpu_workspace(n) >= n * 140
    assert lib.crt_wide_build_gpu_workspace(n, 4) > 0 and lib.crt_wide_build_gpu_workspace(n, 5) == 0
    a8, a64 = lib.crt_wide_build_gpu_async_workspace(n, 4, 8), lib.crt_wide_build_gpu_async_workspace(n, 4, 64)
    assert 0 < a8 < a64 and a64 - a8 >= 56 * 16        # two int64 of level state per extra level
    assert lib.crt_wide_build_gpu_async_workspace(n, 8, 64) > a64    # eight child slots per node
    for bad in ((1, 4, 8), (n, 6, 8), (n, 4, 0), (n, 4, crt.WIDE_GPU_MAX_LEVELS + 1)):
        assert lib.crt_wide_build_gpu_async_workspace(*bad) == 0
    assert lib.crt_wide_build_gpu_async(None, 1, None, 2, None, 1, 4, 8, None, None, None, None, None, 0, None) == -1
    assert b"null" in lib.crt_last_error()
    assert (crt.WIDE_GPU_BAD_LINK, crt.WIDE_GPU_NOT_TREE, crt.WIDE_GPU_BAD_VERTEX, crt.WIDE_GPU_TOO_DEEP) == (1, 2, 4, 8)
    # graph capture needs a created stream and an output
    assert lib.crt_graph_begin(None) == -1
    assert lib.crt_graph_end(None, None) == -1
    assert lib.crt_graph_launch(None, None) == -1
    assert lib.crt_graph_destroy(None) == 0


def test_layout_autotuner_rejects_mismatched_layouts(crt):
    """crt_autotune_batched_layouts checks that every layout holds the same work before any device call."""
    import ctypes as C
    lib = crt.lib()
    sv = crt.SceneView(0, 1, 10, 1, 1, 1, 1, None)
    b = (crt.Batch * 2)(crt.Batch(sv, 0, 0, 0x1000, 0x2000, 64, None), crt.Batch(sv, 0, 0, 0x1100, 0x2000, 64, None))
    best = crt.BatchConfig()
    ms, idx = C.c_float(), C.c_int32()
    rc = lib.crt_autotune_batched_layouts(b, 1, 2, None, 0, 1, None, C.byref(idx), C.byref(best), C.byref(ms))
    assert rc == -1 and b"differs from layout 0" in lib.crt_last_error()
    assert lib.crt_autotune_batched_layouts(b, 0, 2, None, 0, 1, None, C.byref(idx), C.byref(best), C.byref(ms)) == -1
    assert lib.crt_autotune_batched_layouts(None, 1, 1, None, 0, 1, None, None, None, None) == -1

