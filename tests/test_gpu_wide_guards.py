"""Traversal guards of every wide kernel on hand-written BVH4 / BVH8 buffers.

The buffers (oracle_cases.crafted_wide) trip the iteration cap, the stack
capacity, the node-range guard and the speculative walk's record guard; the
host twins' behaviour on them is pinned against the geometry in
test_oracle_wide_guards.py.  Here each kernel's records must equal its twin's
bit for bit: the speculative grid kernel (every preset) the speculative twin
with the preset's schedule, triangle-test counts included; the sequential
kernels (thread per ray at each LDS depth, wave presets, refilling presets)
the sequential twin at their stack capacity (LDS entries + 64); the refilling speculative kernel,
whose wave schedule varies, must set the same guard flags and find the same
t.  Every node and record a walk reads lies inside the buffers: the guards
refuse the bad indices before any read, and the sequential kernels, which read
leaf records unchecked, are not given the bad-record buffer.
"""
import sys

import numpy as np
import pytest

from oracle_cases import WIDE_KINDS, crafted_wide, crafted_wide_rays
from test_gpu_trace import compare

pytestmark = pytest.mark.gpu

WAVE_DEPTH = {3: 16}        # wave presets with 16 LDS stack entries; the other wide presets use 8


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu(crt):
    if crt.device_count() < 1:
        pytest.fail("no HIP device visible")
    return crt


def _upload(gpu, orc, kind, width):
    g, r, nc = crafted_wide(kind, width)
    ents = orc.identity_entity(nc)
    w = gpu.WideBVH(width, g, r, ents, {"node_count": g.size * 4 // width})
    return gpu.DeviceWide(w), w


def _trace(gpu, scene, q, rays, kernel, variant=0, grid=0, lds=0, block=0):
    dr = gpu.DeviceArray.from_host(rays)
    dh = gpu.DeviceArray(gpu.HIT_DTYPE, rays.size)
    tt = gpu.DeviceArray(np.uint32, rays.size)
    dh.fill_bytes(0xAB)
    gpu.trace(scene, q, dr, dh, kernel=kernel, variant=variant, grid=grid, lds_stack=lds, block_size=block,
              tri_tests=tt)
    gpu.lib().crt_device_sync()
    return dh.download(), tt.download()


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("kind", WIDE_KINDS)
def test_speculative_guards(gpu, orc, kind, width):
    dw, w = _upload(gpu, orc, kind, width)
    rays = crafted_wide_rays()
    for q in (0, 1):
        for v in orc.SPEC_PRESETS:
            log(f"[case] speculative grid {kind} w{width} q={q} preset {v}")
            g, gt = _trace(gpu, dw.scene, q, rays, gpu.KERNEL_SPECULATIVE, v)
            o, ot = orc.trace_wide_spec(width, w.groups, w.records, w.entities, q, rays, variant=v, with_tests=True)
            compare(g, o, f"spec guards {kind} w{width} q={q} v={v}")
            assert g.tobytes() == o.tobytes()
            assert np.array_equal(gt, ot)
            log(f"[case] speculative refill {kind} w{width} q={q} preset {v}")
            p, _ = _trace(gpu, dw.scene, q, rays, gpu.KERNEL_SPECULATIVE_REFILL, v)
            assert np.array_equal(p["status"] >> 24, o["status"] >> 24), "guard flags differ"
            assert np.array_equal(p["t"], o["t"]) if q == 0 else np.array_equal(p["t"] > 0, o["t"] > 0)
    log(f"[parity] speculative kernels on crafted {kind} w{width}: equal to the twin")


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("kind", [k for k in WIDE_KINDS if k != "bad_record"])
def test_sequential_guards(gpu, orc, kind, width):
    dw, w = _upload(gpu, orc, kind, width)
    rays = crafted_wide_rays()
    for q in (0, 1):
        twin = {cap: orc.trace_wide(width, w.groups, w.records, w.entities, q, rays, stack_cap=cap)
                for cap in (72, 80, 96)}
        for lds, block in ((0, 0), (8, 0), (16, 0), (32, 128)):      # 0: the defaults, 16 entries, 256 lanes
            log(f"[case] wide thread {kind} w{width} q={q} lds {lds} block {block}")
            g, _ = _trace(gpu, dw.scene, q, rays, gpu.KERNEL_THREAD_PER_RAY, lds=lds, block=block)
            compare(g, twin[(lds or 16) + 64], f"thread guards {kind} w{width} q={q} lds={lds}")
        for v in range(10):
            log(f"[case] wide wave {kind} w{width} q={q} preset {v}")
            g, _ = _trace(gpu, dw.scene, q, rays, gpu.KERNEL_WAVE, v)
            compare(g, twin[WAVE_DEPTH.get(v, 8) + 64], f"wave guards {kind} w{width} q={q} v={v}")
        for v, grid in ((0, 0), (1, 1), (4, 3), (6, 0), (9, 0)):
            log(f"[case] wide refill {kind} w{width} q={q} preset {v} grid {grid}")
            g, _ = _trace(gpu, dw.scene, q, rays, gpu.KERNEL_PERSISTENT, v, grid)
            compare(g, twin[72], f"refill guards {kind} w{width} q={q} v={v}")
    log(f"[parity] sequential wide kernels on crafted {kind} w{width}: equal to the twin")
