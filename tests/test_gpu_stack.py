"""GPU parity of the wave-scheduled and refilling walks of the stack layout.

The reference's stack layout (FlattenStackBVH) is walked by IntersectBVHStack
(Include/TraverseBVHStack.glsl:168-324, occlusion :509-657).  Every preset of
the grid walk (crt_trace kernel 3) and of the refilling walk (kernel 2, over
several grid sizes) must give the host oracle's hit records bit for bit:
closest and any hit, instanced entities with flags, ragged batch sizes, the
63-entry overflow guard (the stack spills past its LDS part on the way), the
triangle-test counters, and a graph-captured refilling launch.
"""
import sys

import numpy as np
import pytest

from oracle_cases import axis_aligned_rays, crafted_stack_chain, random_rays, single_triangle_scene
from test_gpu_trace import compare, device_scene, gpu_trace

pytestmark = pytest.mark.gpu

GRID_PRESETS = range(8)
REFILL_PRESETS = range(8)


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu(crt):
    if crt.device_count() < 1:
        pytest.fail("no HIP device visible")
    return crt


@pytest.fixture(scope="module")
def cloud(gpu):
    import oracle
    from candela_amd import scenes
    b = gpu.build_bvh(1, scenes.icosphere_cloud(count=200, subdivisions=3, seed=17, box=40.0))
    ents = oracle.identity_entity(b.nodes.size)
    s, keep = device_scene(gpu, 1, b.nodes, b.triangles, b.vertices, ents)
    return b, ents, s, keep


@pytest.mark.parametrize("query", [0, 1])
def test_every_preset(gpu, orc, cloud, query):
    b, ents, s, _ = cloud
    rays = np.concatenate([random_rays(40000, [-5, -5, -5], [45, 45, 45], seed=61, tmax=3.0 if query else 1e6),
                           axis_aligned_rays(3000, [0, 0, 0], [40, 40, 40], seed=62)])
    oh = orc.trace(1, b.nodes, b.triangles, b.vertices, ents, query, rays)
    for v in GRID_PRESETS:
        log(f"[case] stack grid preset {v} q={query}")
        compare(gpu_trace(gpu, s, query, rays, kernel=3, variant=v), oh, f"stack grid v={v} q={query}")
    for v in REFILL_PRESETS:
        for grid in ((0, 1, 7, 33) if v in (0, 6) else (0,)):
            log(f"[case] stack refill preset {v} grid {grid} q={query}")
            compare(gpu_trace(gpu, s, query, rays, kernel=2, variant=v, grid=grid), oh,
                    f"stack refill v={v} grid={grid} q={query}")
    log(f"[parity] stack layout q={query}: every grid and refill preset bit-identical to the oracle")


def test_instances_and_flags(gpu, orc):
    from candela_amd import scenes
    r = gpu.RayIntersector(1)
    r.add_object(0, scenes.boxes_scene())
    r.add_object(3, scenes.icosphere_cloud(count=20, subdivisions=2, seed=3, box=16.0))
    rng = np.random.default_rng(7)
    for k in range(6):
        ang = rng.uniform(0, 2 * np.pi)
        m = np.eye(4)
        m[:3, :3] = [[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]]
        m[:3, 3] = rng.uniform(-10, 10, 3)
        r.push_entity(k % 2 * 3, m, translucency=0.5 if k == 2 else 0.0)
    r.buffer_entities()
    r.buffer_data()
    hb = r.host_buffers()
    rays = random_rays(20000, [-25, -5, -25], [25, 20, 25], seed=63, toward=[0, 2, 0])
    for flags in (0, gpu.FLAG_IGNORE_TRANSPARENT, gpu.FLAG_TRI0_IS_MISS):
        for query in (0, 1):
            oh = orc.trace(1, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], query, rays, flags=flags)
            for kernel, variant in ((3, 0), (3, 7), (2, 0), (2, 6)):
                gh = gpu_trace(gpu, r.scene(), query, rays, flags=flags, kernel=kernel, variant=variant)
                compare(gh, oh, f"stack instances k={kernel}:{variant} flags={flags} q={query}")
    log("[parity] stack layout instances and flags bit-identical to the oracle")


def test_overflow_guard_and_spill(gpu, orc):
    import oracle
    v, tris = single_triangle_scene()
    ray = random_rays(64, [0.1, 0.1, 1], [0.3, 0.3, 1], seed=0)
    ray["dir"] = [0, 0, -1]
    for depth, flag in ((70, 1 << 25), (40, 0)):   # 63 pushes overflow; 39 fit (spilling past 8 or 16 LDS entries)
        nodes = crafted_stack_chain(depth)
        ents = oracle.identity_entity(nodes.size)
        s, keep = device_scene(gpu, 1, nodes, tris, v, ents)
        oh = orc.trace(1, nodes, tris, v, ents, 0, ray)
        assert (oh["status"] & 0xFF000000 == flag).all()
        for kernel, variant in [(3, p) for p in GRID_PRESETS] + [(2, p) for p in REFILL_PRESETS]:
            compare(gpu_trace(gpu, s, 0, ray, kernel=kernel, variant=variant), oh,
                    f"chain {depth} k={kernel}:{variant}")
    log("[parity] stack chain 70 (overflow) and 40 (spill) on every preset")


def test_ragged_batches_and_counters(gpu, orc, cloud):
    b, ents, s, _ = cloud
    for n in (0, 1, 63, 65, 257, 1000, 4097):
        rays = random_rays(max(n, 1), [-5, -5, -5], [45, 45, 45], seed=70 + n)[:n]
        dr = gpu.DeviceArray(gpu.RAY_DTYPE, max(n, 1))
        if n:
            dr.upload(rays)
        dh = gpu.DeviceArray(gpu.HIT_DTYPE, max(n, 1))
        for kernel in (2, 3):
            dh.fill_bytes(0x7F)
            tt = gpu.DeviceArray(np.uint32, max(n, 1))
            gpu.trace(s, 0, dr, dh, n=n, kernel=kernel, tri_tests=tt)
            gpu.lib().crt_device_sync()
            out = dh.download()
            if n == 0:
                assert out.view(np.uint8).tolist() == [0x7F] * 32
                continue
            oh, otests = orc.trace(1, b.nodes, b.triangles, b.vertices, ents, 0, rays, with_tests=True)
            compare(out[:n], oh, f"stack n={n} kernel={kernel}")
            assert np.array_equal(tt.download()[:n], otests), f"triangle tests differ (n={n}, kernel={kernel})"
    log("[parity] stack ragged batches and triangle-test counters")


def test_refill_in_a_graph(gpu, orc, cloud):
    b, ents, s, _ = cloud
    rays = random_rays(30000, [-5, -5, -5], [45, 45, 45], seed=80)
    oh = orc.trace(1, b.nodes, b.triangles, b.vertices, ents, 0, rays)
    dr = gpu.DeviceArray.from_host(rays)
    dh = gpu.DeviceArray(gpu.HIT_DTYPE, rays.size)
    st = gpu.Stream()
    q = gpu.queue_block()
    g = gpu.Graph(st)
    with g:
        gpu.trace(s, 0, dr, dh, kernel=2, variant=0, stream=st, queue=q)
    for rep in range(3):
        dh.fill_bytes(0xFF, stream=st)
        g.launch()
        st.sync()
        compare(dh.download(), oh, f"stack refill graph replay {rep}")
    g.close()
    st.close()
