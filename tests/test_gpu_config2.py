"""GPU parity on the incoherent GI / RTAO workload at its full size.

2,000 subdivision-3 icospheres (2,560,000 triangles, seed 2), one SAH BVH from
the CPU constructor, a 1920x1080 G-buffer of primary closest hits, and per hit
pixel one cosine RTAO any-hit ray (tmax 2) and one cosine GI closest-hit ray
(the rays DiffuseTrace.glsl:445-446 spawns before IntersectRay at :484 and
:494, with the 1e-4 offset and PCG directions of this workload).  Every ray is
traced on the GPU and compared with the host oracle over the same buffers:

- BVH2 kernels (stackless layout): bit-identical to the reference walk;
- the sequential BVH4 walks (refilling walk for GI, grid walk in slot order
  for RTAO): bit-identical to the host twin of the wide walk, and GI
  triangle ids agree with the reference walk on at least 99.999 % of rays,
  every mismatch at an exactly equal t (the same distance reached through
  another triangle of a shared edge), RTAO occlusion on every ray;
- the speculative BVH4 kernels the benchmark picks (refilling speculative
  walk for GI, speculative grid walk for RTAO, over the greedy and the SAH
  collapse): bit-identical to the host twin run with each preset's schedule
  (a ray's record does not depend on which rays share its wave, so this
  holds for the refilling launch as well); GI nearest t equal to the
  sequential wide walk's on every ray, and the same bar against the
  reference walk as above;
- the stack layout (FlattenStackBVH of the same meshes): bit-identical to the
  reference's stack walk.
"""
import sys

import numpy as np
import pytest

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
def work(gpu):
    from candela_amd import workloads
    log("[case] config 2: building 2,000 icospheres and the 1920x1080 G-buffer")
    w = workloads.incoherent_cloud(gpu)
    log(f"[case] config 2: {w.triangles} triangles, {w.ao.size} AO + {w.gi.size} GI rays")
    assert w.triangles == 2_560_000 and w.ao.size == w.gi.size > 1_000_000
    return w


@pytest.fixture(scope="module")
def ref(orc, work):
    """The reference walk (stackless, BVH2) over every GI and AO ray."""
    hb = work.host
    g = orc.trace(0, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], 0, work.gi)
    a = orc.trace(0, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], 1, work.ao)
    return g, a


def _trace(gpu, scene, query, rays, kernel, variant, flags=0):
    d = gpu.DeviceArray.from_host(rays)
    h = gpu.DeviceArray(gpu.HIT_DTYPE, rays.size)
    h.fill_bytes(0xAB)
    gpu.trace(scene, query, d, h, kernel=kernel, variant=variant, flags=flags)
    gpu.lib().crt_device_sync()
    return h.download()


def _differs(a, b):
    return int(np.count_nonzero((a.view(np.uint32).reshape(-1, 8) != b.view(np.uint32).reshape(-1, 8)).any(1)))


@pytest.mark.parametrize("kernel,variant", [(3, 1), (2, 0), (1, 0)])
def test_config2_bvh2_bit_identical(gpu, work, ref, kernel, variant):
    log(f"[case] config 2 BVH2 kernel {kernel}:{variant}")
    g = _trace(gpu, work.scene, 0, work.gi, kernel, variant)
    a = _trace(gpu, work.scene, 1, work.ao, kernel, variant)
    assert _differs(g, ref[0]) == 0, "GI records differ from the reference walk"
    assert _differs(a, ref[1]) == 0, "AO records differ from the reference walk"
    log(f"[parity] config 2 BVH2 {kernel}:{variant}: {g.size + a.size} rays bit-identical to the reference walk")


def test_config2_bvh4_benchmark_choice(gpu, orc, work, ref):
    log("[case] config 2 BVH4: GI refill preset 8, AO grid preset 7 in slot order")
    dw = gpu.upload_wide(gpu.FORMAT_STACKLESS, work.host, 4, 2)
    w = dw.host
    g = _trace(gpu, dw.scene, 0, work.gi, gpu.KERNEL_PERSISTENT, 8)
    a = _trace(gpu, dw.scene, 1, work.ao, gpu.KERNEL_WAVE, 7, flags=gpu.FLAG_SLOT_ORDER)
    tg = orc.trace_wide(w.width, w.groups, w.records, w.entities, 0, work.gi)
    ta = orc.trace_wide(w.width, w.groups, w.records, w.entities, 1, work.ao, flags=gpu.FLAG_SLOT_ORDER)
    assert _differs(g, tg) == 0, "GI records differ from the host twin"
    assert _differs(a, ta) == 0, "AO records differ from the host twin"
    rg, ra = ref
    miss = g["tri"] != rg["tri"]
    agree = 1.0 - miss.mean()
    log(f"[parity] config 2 BVH4: exact vs twin; GI id agreement vs reference walk {agree:.7f} "
        f"({int(miss.sum())} mismatches, {int((miss & (g['t'] == rg['t'])).sum())} at equal t)")
    assert agree >= 0.99999
    assert (g["t"][miss] == rg["t"][miss]).all(), "an id mismatch at a different distance"
    assert np.array_equal(g["t"] > 0, rg["t"] > 0)
    assert np.array_equal(a["t"] > 0, ra["t"] > 0), "RTAO occlusion differs from the reference walk"
    # barycentrics within 1e-5 of the reference walk where the triangle agrees
    same = ~miss & (g["t"] > 0)
    for k in ("u", "v", "w"):
        assert np.abs(g[k][same] - rg[k][same]).max() <= 1e-5


def test_config2_bvh4_speculative_choice(gpu, orc, work, ref):
    log("[case] config 2 BVH4: GI refilling speculative preset 4, AO speculative grid preset 5")
    dw = gpu.upload_wide(gpu.FORMAT_STACKLESS, work.host, 4, 2)
    w = dw.host
    g = _trace(gpu, dw.scene, 0, work.gi, gpu.KERNEL_SPECULATIVE_REFILL, 4)
    a = _trace(gpu, dw.scene, 1, work.ao, gpu.KERNEL_SPECULATIVE, 5)
    ta = orc.trace_wide_spec(w.width, w.groups, w.records, w.entities, 1, work.ao, variant=5)
    assert _differs(a, ta) == 0, "AO records differ from the twin of the speculative grid walk"
    tw = orc.trace_wide_spec(w.width, w.groups, w.records, w.entities, 0, work.gi, schedule=orc.SPEC_P_PRESETS[4])
    assert _differs(g, tw) == 0, "GI records differ from the twin run with the refilling preset's schedule"
    tg = orc.trace_wide(w.width, w.groups, w.records, w.entities, 0, work.gi)
    assert np.array_equal(g["t"], tg["t"]), "GI nearest t differs from the sequential wide walk"
    assert (g["status"] >> 24 == 0).all()
    rg, ra = ref
    miss = g["tri"] != rg["tri"]
    agree = 1.0 - miss.mean()
    log(f"[parity] config 2 speculative BVH4: AO exact vs twin; GI id agreement vs reference walk {agree:.7f} "
        f"({int(miss.sum())} mismatches, {int((miss & (g['t'] == rg['t'])).sum())} at equal t)")
    assert agree >= 0.99999
    assert (g["t"][miss] == rg["t"][miss]).all(), "an id mismatch at a different distance"
    assert np.array_equal(g["t"] > 0, rg["t"] > 0)
    assert np.array_equal(a["t"] > 0, ra["t"] > 0), "RTAO occlusion differs from the reference walk"
    same = ~miss & (g["t"] > 0)
    for k in ("u", "v", "w"):
        assert np.abs(g[k][same] - rg[k][same]).max() <= 1e-5


def test_config2_sah_bvh4_choice(gpu, orc, work, ref):
    """The headline choice w4s2:5:2+ao:4:5: the SAH-collapsed BVH4 with 2-triangle leaves (tri_cost 0.3), GI
    refilling speculative preset 2, RTAO speculative grid preset 5 -- every record of both launches bit-identical
    to the twin with each preset's schedule."""
    log("[case] config 2 SAH BVH4: GI refilling speculative preset 2, AO speculative grid preset 5")
    dw = gpu.upload_wide(gpu.FORMAT_STACKLESS, work.host, 4, 2, None, None, gpu.COLLAPSE_SAH, 0.0, 0.3)
    w = dw.host
    g = _trace(gpu, dw.scene, 0, work.gi, gpu.KERNEL_SPECULATIVE_REFILL, 2)
    a = _trace(gpu, dw.scene, 1, work.ao, gpu.KERNEL_SPECULATIVE, 5)
    tg = orc.trace_wide_spec(w.width, w.groups, w.records, w.entities, 0, work.gi, schedule=orc.SPEC_P_PRESETS[2])
    ta = orc.trace_wide_spec(w.width, w.groups, w.records, w.entities, 1, work.ao, variant=5)
    assert _differs(g, tg) == 0, "GI records differ from the twin"
    assert _differs(a, ta) == 0, "AO records differ from the twin"
    rg, ra = ref
    miss = g["tri"] != rg["tri"]
    log(f"[parity] config 2 SAH BVH4 (w4s2:5:2+ao:4:5): GI and AO bit-identical to the twin; GI id agreement vs "
        f"reference walk {1.0 - miss.mean():.7f}")
    assert (g["t"][miss] == rg["t"][miss]).all(), "an id mismatch at a different distance"
    assert np.array_equal(g["t"] > 0, rg["t"] > 0)
    assert np.array_equal(a["t"] > 0, ra["t"] > 0), "RTAO occlusion differs from the reference walk"


def test_config2_stack_layout(gpu, orc, work):
    from candela_amd import scenes
    log("[case] config 2 stack layout: FlattenStackBVH of the same meshes")
    r = gpu.RayIntersector(gpu.FORMAT_STACK)
    r.add_object(0, scenes.icosphere_cloud(count=2000, subdivisions=3, seed=2))
    r.push_entity(0)
    r.buffer_entities()
    r.buffer_data()
    hb = r.host_buffers()
    sc = r.scene()
    og = orc.trace(1, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], 0, work.gi)
    oa = orc.trace(1, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], 1, work.ao)
    for kernel, variant in [(1, 0), (3, 0), (3, 6), (2, 0), (2, 6)]:
        g = _trace(gpu, sc, 0, work.gi, kernel, variant)
        a = _trace(gpu, sc, 1, work.ao, kernel, variant)
        assert _differs(g, og) == 0, f"stack GI records differ ({kernel}:{variant})"
        assert _differs(a, oa) == 0, f"stack AO records differ ({kernel}:{variant})"
        log(f"[parity] config 2 stack layout {kernel}:{variant}: bit-identical to the reference stack walk")
