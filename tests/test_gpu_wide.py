"""GPU parity of the wide (BVH4 / BVH8) walk.

The GPU wide walk and the host walk of the same wide layout (oracle_trace_wide)
do the same FP32 operations in the same order, so every hit record and every
triangle-test count must be bit-identical.  Against the reference semantics,
the nearest distance must equal the BVH2 stack walk's (which is the
brute-force nearest hit) on every ray, with the triangle equal up to exact
distance ties.
"""
import sys

import numpy as np
import pytest

from oracle_cases import axis_aligned_rays, random_rays
from test_gpu_trace import compare, gpu_trace

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu(crt):
    if crt.device_count() < 1:
        pytest.fail("no HIP device visible")
    return crt


def _wide_context(crt, fmt, meshes, width, max_leaf):
    r = crt.RayIntersector(fmt)
    r.add_object(0, meshes)
    r.push_entity(0)
    r.buffer_entities()
    r.buffer_data()
    r.build_wide(width, max_leaf)
    return r


def _host_wide(crt, r, width, max_leaf):
    hb = r.host_buffers()
    return hb, crt.wide_build(r.format, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], width,
                              max_leaf)


@pytest.mark.parametrize("width,max_leaf", [(4, 2), (8, 2), (4, 4), (8, 6)])
def test_wide_boxes(gpu, orc, width, max_leaf):
    from candela_amd import scenes
    log(f"[case] wide boxes width={width} max_leaf={max_leaf}")
    r = _wide_context(gpu, 0, scenes.boxes_scene(), width, max_leaf)
    hb, w = _host_wide(gpu, r, width, max_leaf)
    rays = np.concatenate([random_rays(20000, [-4, 0.2, -4], [20, 8, 20], seed=3),
                           random_rays(10000, [-10, 5, -10], [-5, 9, -5], seed=4, toward=[8, 1, 8]),
                           axis_aligned_rays(4000, [-2, 0.1, -2], [18, 5, 18], seed=5)])
    scene = r.wide_scene()
    assert scene.format == (gpu.FORMAT_WIDE8 if width == 8 else gpu.FORMAT_WIDE4)
    for q in (0, 1):
        g = gpu_trace(gpu, scene, q, rays)
        o = orc.trace_wide(width, w.groups, w.records, w.entities, q, rays)
        compare(g, o, f"wide{width} leaf{max_leaf} boxes q={q}")
    g = gpu_trace(gpu, scene, 0, rays)
    st = gpu.RayIntersector(1)
    st.add_object(0, scenes.boxes_scene())
    sb = st.host_buffers()
    import oracle
    ref = orc.trace(1, sb["nodes"], sb["triangles"], sb["vertices"], oracle.identity_entity(sb["nodes"].size), 0,
                    rays)
    assert np.array_equal(g["t"], ref["t"])
    log(f"[parity] wide{width} vs BVH2 stack walk: t identical on {rays.size} rays, "
        f"same triangle on {np.mean(g['tri'] == ref['tri']):.5f} (the rest are exact distance ties)")


@pytest.mark.parametrize("src", [0, 1])
@pytest.mark.parametrize("width", [4, 8])
def test_wide_cloud_incoherent(gpu, orc, src, width):
    """A 200-sphere cloud with incoherent rays, from either BVH2 source format; counters too."""
    from candela_amd import scenes
    log(f"[case] wide cloud src={src} width={width}")
    meshes = scenes.icosphere_cloud(count=200, subdivisions=3, seed=17, box=40.0)
    r = _wide_context(gpu, src, meshes, width, 2)
    hb, w = _host_wide(gpu, r, width, 2)
    rays = random_rays(100000, [-2, -2, -2], [42, 42, 42], seed=18)
    scene = r.wide_scene()
    for q in (0, 1):
        dr = gpu.DeviceArray.from_host(rays)
        dh = gpu.DeviceArray(gpu.HIT_DTYPE, rays.size)
        dt = gpu.DeviceArray(np.uint32, rays.size)
        gpu.trace(scene, q, dr, dh, tri_tests=dt)
        gpu.lib().crt_device_sync()
        o, ot = orc.trace_wide(width, w.groups, w.records, w.entities, q, rays, with_tests=True)
        compare(dh.download(), o, f"wide{width} src{src} cloud q={q}")
        assert np.array_equal(dt.download(), ot)


@pytest.mark.parametrize("width,max_leaf", [(4, 2), (4, 4), (8, 2)])
def test_wide_sah_context(gpu, orc, width, max_leaf):
    """crt_context_build_wide_ex with the SAH collapse: the context's layout walks (wave kernel, speculative grid
    and refilling kernels) against the host twins over crt_wide_build_ex's layout of the same buffers."""
    from candela_amd import scenes
    log(f"[case] SAH context width={width} max_leaf={max_leaf}")
    r = gpu.RayIntersector(0)
    r.add_object(0, scenes.icosphere_cloud(count=150, subdivisions=3, seed=29, box=35.0))
    r.push_entity(0)
    r.buffer_entities()
    r.buffer_data()
    r.build_wide(width, max_leaf, gpu.COLLAPSE_SAH, 0.0, 0.3)
    hb = r.host_buffers()
    w = gpu.wide_build(0, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], width, max_leaf,
                       gpu.COLLAPSE_SAH, 0.0, 0.3)
    rays = np.concatenate([random_rays(40000, [-2, -2, -2], [37, 37, 37], seed=30),
                           axis_aligned_rays(3000, [0, 0, 0], [35, 35, 35], seed=31)])
    scene = r.wide_scene()
    for q in (0, 1):
        o = orc.trace_wide(width, w.groups, w.records, w.entities, q, rays, stack_cap=8 + 64)
        compare(gpu_trace(gpu, scene, q, rays, kernel=3, variant=0), o, f"SAH w{width} l{max_leaf} wave q={q}")
        os_ = orc.trace_wide_spec(width, w.groups, w.records, w.entities, q, rays, variant=1)
        compare(gpu_trace(gpu, scene, q, rays, kernel=gpu.KERNEL_SPECULATIVE, variant=1), os_,
                f"SAH w{width} l{max_leaf} speculative q={q}")
        g = gpu_trace(gpu, scene, q, rays, kernel=gpu.KERNEL_SPECULATIVE_REFILL, variant=4)
        assert np.array_equal(g["t"] > 0, o["t"] > 0)
        if q == 0:
            assert np.array_equal(g["t"], o["t"])
    log(f"[parity] SAH context layout width {width} leaves {max_leaf}: equal to the twins")


@pytest.mark.parametrize("width", [4, 8])
def test_wide_wave_presets(gpu, orc, width):
    """The wave-scheduled and the refilling wide walks (every preset, including the occupancy-hinted builds
    6..9, several grids) against the host twin."""
    from candela_amd import scenes
    log(f"[case] wide wave presets width={width}")
    meshes = scenes.icosphere_cloud(count=150, subdivisions=3, seed=23, box=35.0)
    r = _wide_context(gpu, 0, meshes, width, 2)
    hb, w = _host_wide(gpu, r, width, 2)
    rays = np.concatenate([random_rays(50000, [-2, -2, -2], [37, 37, 37], seed=24),
                           random_rays(20000, [-2, -2, -2], [37, 37, 37], seed=25, tmax=2.0)])
    scene = r.wide_scene()
    for q in (0, 1):
        o = orc.trace_wide(width, w.groups, w.records, w.entities, q, rays, stack_cap=8 + 64)
        o16 = orc.trace_wide(width, w.groups, w.records, w.entities, q, rays, stack_cap=16 + 64)
        assert np.array_equal(o, o16)                   # no ray comes near the stack capacity
        compare(gpu_trace(gpu, scene, q, rays, kernel=1), o, f"wide{width} thread q={q}")
        for v in range(10):
            compare(gpu_trace(gpu, scene, q, rays, kernel=3, variant=v), o, f"wide{width} wave v{v} q={q}")
        for v, grid in ((0, 0), (1, 7), (2, 0), (3, 64), (4, 0), (5, 1), (6, 0), (7, 3), (8, 0), (9, 0)):
            compare(gpu_trace(gpu, scene, q, rays, kernel=2, variant=v, grid=grid), o,
                    f"wide{width} refill v{v} grid={grid} q={q}")


@pytest.mark.parametrize("width", [4, 8])
def test_wide_slot_order(gpu, orc, width):
    """Any-hit walks with CRT_FLAG_SLOT_ORDER (children in slot order) on every wide kernel and preset against
    the host twin with the same flag; closest-hit launches ignore the flag."""
    from candela_amd import scenes
    log(f"[case] wide slot order width={width}")
    meshes = scenes.icosphere_cloud(count=150, subdivisions=3, seed=33, box=35.0)
    r = _wide_context(gpu, 0, meshes, width, 2)
    hb, w = _host_wide(gpu, r, width, 2)
    rays = np.concatenate([random_rays(40000, [-2, -2, -2], [37, 37, 37], seed=34),
                           random_rays(20000, [-2, -2, -2], [37, 37, 37], seed=35, tmax=2.0)])
    scene = r.wide_scene()
    f = gpu.FLAG_SLOT_ORDER
    o = orc.trace_wide(width, w.groups, w.records, w.entities, 1, rays, flags=f, stack_cap=8 + 64)
    assert not np.array_equal(o, orc.trace_wide(width, w.groups, w.records, w.entities, 1, rays, stack_cap=8 + 64))
    compare(gpu_trace(gpu, scene, 1, rays, kernel=1, flags=f), o, f"wide{width} thread slot order")
    for v in range(10):
        compare(gpu_trace(gpu, scene, 1, rays, kernel=3, variant=v, flags=f), o, f"wide{width} wave v{v} slot order")
    for v, grid in ((0, 0), (2, 5), (6, 0), (8, 0), (9, 0)):
        compare(gpu_trace(gpu, scene, 1, rays, kernel=2, variant=v, grid=grid, flags=f), o,
                f"wide{width} refill v{v} grid={grid} slot order")
    oc = orc.trace_wide(width, w.groups, w.records, w.entities, 0, rays, stack_cap=8 + 64)
    compare(gpu_trace(gpu, scene, 0, rays, kernel=2, variant=8, flags=f), oc, f"wide{width} closest ignores it")


@pytest.mark.parametrize("block,lds", [(64, 8), (256, 16), (128, 32), (512, 8)])
def test_wide_launch_shapes(gpu, orc, block, lds):
    from candela_amd import scenes
    log(f"[case] wide launch block={block} lds={lds}")
    meshes = scenes.icosphere_cloud(count=60, subdivisions=2, seed=19, box=20.0)
    r = _wide_context(gpu, 0, meshes, 8, 2)
    hb, w = _host_wide(gpu, r, 8, 2)
    rays = random_rays(30001, [-2, -2, -2], [22, 22, 22], seed=20)
    g = gpu_trace(gpu, r.wide_scene(), 0, rays, block=block, lds=lds, kernel=1)
    o = orc.trace_wide(8, w.groups, w.records, w.entities, 0, rays, stack_cap=lds + 64)
    compare(g, o, f"wide8 block={block} lds={lds}")


def test_wide_instances_and_flags(gpu, orc):
    from candela_amd import scenes
    log("[case] wide instances")
    r = gpu.RayIntersector(0)
    r.add_object(0, scenes.boxes_scene())
    r.add_object(1, scenes.icosphere_cloud(count=20, subdivisions=2, seed=5, box=10.0))
    up = np.eye(4, dtype=np.float32)
    up[1, 3] = 6.0
    r.push_entity(0)
    r.push_entity(1, up, translucency=0.5)
    r.push_entity(0, up)
    r.buffer_entities()
    r.buffer_data()
    r.build_wide(4, 2)
    hb = r.host_buffers()
    w = gpu.wide_build(0, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], 4, 2)
    rays = random_rays(40000, [-2, 0.2, -2], [18, 14, 18], seed=21)
    for q in (0, 1):
        for flags in (0, 1, 2, gpu.FLAG_SLOT_ORDER, gpu.FLAG_SLOT_ORDER | 1):
            g = gpu_trace(gpu, r.wide_scene(), q, rays, flags=flags)
            o = orc.trace_wide(4, w.groups, w.records, w.entities, q, rays, flags=flags)
            compare(g, o, f"wide4 instances q={q} flags={flags}")


def test_wide_errors(gpu):
    from candela_amd import scenes
    r = _wide_context(gpu, 0, scenes.boxes_scene(), 4, 2)
    rays = random_rays(10, [0, 0, 0], [1, 1, 1], seed=0)
    with pytest.raises(gpu.CrtError):
        gpu_trace(gpu, r.wide_scene(), 0, rays, kernel=3, variant=10)
    with pytest.raises(gpu.CrtError):
        gpu_trace(gpu, r.wide_scene(), 0, rays, kernel=2, variant=10)
    with pytest.raises(gpu.CrtError):
        gpu_trace(gpu, r.wide_scene(), 0, rays, lds=64)            # wide: 8, 16 or 32
    with pytest.raises(gpu.CrtError):
        gpu_trace(gpu, r.wide_scene(), 0, rays, block=1024, lds=16)  # (2 * 16 + 8) * 1024 * 4 > 64 KiB
    r.buffer_entities()                                            # makes the wide layout stale
    with pytest.raises(gpu.CrtError):
        r.wide_scene()
