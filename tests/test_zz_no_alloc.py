"""crt_trace_noalloc against crt_trace on the same inputs, through the C ABI.

crt_trace_noalloc allocates nothing: the persistent kernels' work-queue
counters live in a static device block per GPU instead of a pool block.  Its
records must equal crt_trace's bit for bit, on two calls in a row of which the
second traces fewer rays (the block keeps the first call's counters, which the
launch must clear), for the persistent kernels over the stackless, stack and
BVH4 layouts -- the refilling speculative kernels included, whose records do
not depend on which rays share a wave -- and for a kernel that needs no
counters.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle_cases import random_rays

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu(crt):
    if crt.device_count() < 1:
        pytest.fail("no HIP device visible")
    return crt


@pytest.fixture(scope="module")
def scenes(gpu):
    from candela_amd import scenes as sc
    r = gpu.RayIntersector(gpu.FORMAT_STACKLESS)
    r.add_object(0, sc.icosphere_cloud(count=300, subdivisions=3, seed=21, box=40.0))
    r.push_entity(0)
    r.buffer_entities()
    r.buffer_data()
    wide = gpu.upload_wide(gpu.FORMAT_STACKLESS, r.host_buffers(), 4, 2)
    rs = gpu.RayIntersector(gpu.FORMAT_STACK)
    rs.add_object(0, sc.icosphere_cloud(count=300, subdivisions=3, seed=21, box=40.0))
    rs.push_entity(0)
    rs.buffer_entities()
    rs.buffer_data()
    return {"stackless": (r, r.scene()), "bvh4": (wide, wide.scene), "stack": (rs, rs.scene())}


def _call(gpu, fn, scene, q, rays, n, kernel, variant, tri_tests):
    hits = gpu.DeviceArray(gpu.HIT_DTYPE, rays.count)
    hits.fill_bytes(0xCD)
    cfg = gpu.TraceConfig(0, 0, kernel, variant, 0, 0, tri_tests.ptr, None)
    rc = fn(C.byref(scene), q, 0, rays.ptr, hits.ptr, n, C.byref(cfg), None)
    assert rc == 0, gpu.lib().crt_last_error()
    gpu.lib().crt_device_sync()
    return hits.download()


@pytest.mark.parametrize("layout,kernel,variant", [("stackless", 2, 0), ("stackless", 2, 3), ("bvh4", 2, 8),
                                                   ("bvh4", 2, 0), ("bvh4", 3, 7), ("bvh4", 5, 2),
                                                   ("stackless", 5, 3), ("stack", 5, 2), ("stack", 2, 4)])
def test_noalloc_equals_crt_trace(gpu, scenes, layout, kernel, variant):
    _, scene = scenes[layout]
    lib = gpu.lib()
    for q in (0, 1):
        rays = random_rays(40000, [-5, -5, -5], [45, 45, 45], seed=200 + q, tmax=3.0 if q else 1e6)
        d_rays = gpu.DeviceArray.from_host(rays)
        for n in (rays.size, 1237):   # the second call traces fewer rays over the same static counters
            log(f"[case] noalloc {layout} kernel {kernel}:{variant} q={q} n={n}")
            tt_a = gpu.DeviceArray(np.uint32, rays.size)
            tt_b = gpu.DeviceArray(np.uint32, rays.size)
            a = _call(gpu, lib.crt_trace_noalloc, scene, q, d_rays, n, kernel, variant, tt_a)
            b = _call(gpu, lib.crt_trace, scene, q, d_rays, n, kernel, variant, tt_b)
            assert a[:n].tobytes() == b[:n].tobytes(), f"{layout} {kernel}:{variant} q={q} n={n}: records differ"
            assert np.array_equal(tt_a.download()[:n], tt_b.download()[:n])
            # nothing past n is written
            assert a.view(np.uint8)[n * 32:].tolist() == [0xCD] * ((rays.size - n) * 32)
            assert (a["t"][:n] > 0).any()
    log(f"[parity] crt_trace_noalloc equals crt_trace: {layout} kernel {kernel}:{variant}")


def test_noalloc_rejects_a_queue_and_too_many_rays(gpu, scenes):
    _, scene = scenes["bvh4"]
    lib = gpu.lib()
    rays = gpu.DeviceArray.from_host(random_rays(64, [0, 0, 0], [1, 1, 1], seed=3))
    hits = gpu.DeviceArray(gpu.HIT_DTYPE, 64)
    q = gpu.queue_block()
    cfg = gpu.TraceConfig(0, 0, 2, 0, 0, 0, None, q.ptr)
    assert lib.crt_trace_noalloc(C.byref(scene), 0, 0, rays.ptr, hits.ptr, 64, C.byref(cfg), None) == -1
    assert b"queue" in lib.crt_last_error()
    cfg = gpu.TraceConfig(0, 0, 2, 0, 0, 0, None, None)
    assert lib.crt_trace_noalloc(C.byref(scene), 0, 0, rays.ptr, hits.ptr, 1 << 31, C.byref(cfg),
                                 None) == -1
    assert b"CRT_NOALLOC_MAX_RAYS" in lib.crt_last_error()
