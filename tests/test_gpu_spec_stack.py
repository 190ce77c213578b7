"""GPU parity of the speculative walk of the reference's stack layout (crt_trace kernels 4 and 5 over
CRT_FORMAT_STACK).

Grid launch (kernel 4): 64 consecutive rays share a wave, so the schedule is fixed and every record -- t,
barycentrics, triangle, mesh, entity, visits, status and the triangle-test count -- must equal the host twin that
replays it (oracle_trace_stack_spec), for every preset and both queries, on a cloud, on the boxes scene, on instanced
entities with the scene flags and on the crafted guard buffers (the overflow chain reaches the stack entries kept in
per-lane scratch).
Refilling launch (kernel 5): which rays share a wave varies from run to run, but a ray's record does not depend on
it (tests/test_oracle_spec_regroup.py), so every record must equal the twin run with the refilling preset's schedule
(oracle.SBVH2_P_PRESETS); the nearest t also equals the reference walk's (oracle_trace), the triangle up to exactly
equal t, and an any-hit ray reports the reference walk's first accepted triangle.
"""
import sys

import numpy as np
import pytest

from oracle_cases import axis_aligned_rays, crafted_stack, crafted_stack_chain, random_rays, single_triangle_scene
from test_gpu_trace import compare, device_scene

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu(crt):
    if crt.device_count() < 1:
        pytest.fail("no HIP device visible")
    return crt


def _context(gpu, meshes):
    r = gpu.RayIntersector(gpu.FORMAT_STACK)
    r.add_object(0, meshes)
    r.push_entity(0)
    r.buffer_entities()
    r.buffer_data()
    return r, r.host_buffers()


@pytest.fixture(scope="module")
def cloud(gpu):
    from candela_amd import scenes
    return _context(gpu, scenes.icosphere_cloud(count=300, subdivisions=3, seed=12, box=40.0))


def _trace(gpu, scene, q, rays, kernel, variant, grid=0, flags=0, n=None):
    dr = gpu.DeviceArray.from_host(rays)
    dh = gpu.DeviceArray(gpu.HIT_DTYPE, max(1, rays.size))
    tt = gpu.DeviceArray(np.uint32, max(1, rays.size))
    dh.fill_bytes(0xAB)
    gpu.trace(scene, q, dr, dh, n=n, kernel=kernel, variant=variant, grid=grid, flags=flags, tri_tests=tt)
    gpu.lib().crt_device_sync()
    return dh.download(), tt.download()


def _rays(q):
    return np.concatenate([random_rays(40000, [-5, -5, -5], [45, 45, 45], seed=101, tmax=3.0 if q else 1e6),
                           axis_aligned_rays(3000, [0, 0, 0], [40, 40, 40], seed=102)])


def _twin_equal(gpu, orc, scene, hb, q, rays, v, flags=0, what=""):
    g, gt = _trace(gpu, scene, q, rays, gpu.KERNEL_SPECULATIVE, v, flags=flags)
    o, ot = orc.trace_stack_spec(hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], q, rays, variant=v,
                                 flags=flags, with_tests=True)
    compare(g, o, f"spec stack grid {what} q={q} v={v} flags={flags}")
    assert g.tobytes() == o.tobytes()
    assert np.array_equal(gt, ot), f"triangle tests differ ({what} q={q} v={v})"


@pytest.mark.parametrize("q", [0, 1])
def test_grid_equals_twin(gpu, orc, cloud, q):
    r, hb = cloud
    rays = _rays(q)
    for v in orc.SBVH2_PRESETS:
        log(f"[case] speculative stack grid cloud q={q} preset {v}")
        _twin_equal(gpu, orc, r.scene(), hb, q, rays, v, what="cloud")
    log(f"[parity] speculative stack grid q={q}: every preset bit-identical to the twin")


def test_refill_equals_reference(gpu, orc, cloud):
    r, hb = cloud
    for q in (0, 1):
        rays = _rays(q)
        ref = orc.trace(1, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], q, rays)
        for v in orc.SBVH2_PRESETS:
            for grid in ((0, 1, 7) if v == 1 else (0,)):
                log(f"[case] speculative stack refill q={q} preset {v} grid {grid}")
                g, gt = _trace(gpu, r.scene(), q, rays, gpu.KERNEL_SPECULATIVE_REFILL, v, grid=grid)
                o, ot = orc.trace_stack_spec(hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], q, rays,
                                 schedule=orc.SBVH2_P_PRESETS[v], with_tests=True)
                compare(g, o, f"refill q={q} v={v} grid={grid}")
                assert g.tobytes() == o.tobytes() and np.array_equal(gt, ot), "not bit-identical to the twin"
                assert np.array_equal(g["t"], ref["t"]), "t differs"
                assert (g["status"] & 0xFF000000 == 0).all()
                if q == 0:
                    same = g["tri"] == ref["tri"]
                    assert (same | (g["t"] == ref["t"])).all() and same.mean() > 0.999
                    for k in ("u", "v", "w"):
                        assert np.array_equal(g[k][same], ref[k][same])
                else:
                    assert np.array_equal(g["tri"], ref["tri"]), "any hit: not the reference walk's first triangle"
    log("[parity] speculative stack refill: every record bit-identical to the twin; nearest t, any-hit triangle and "
        "occlusion equal the reference walk")


def test_boxes_instances_flags_and_ragged(gpu, orc):
    from candela_amd import scenes
    r, hb = _context(gpu, scenes.boxes_scene())
    rays = random_rays(20000, [-4, 0.2, -4], [20, 8, 20], seed=3)
    for q in (0, 1):
        for v in (0, 3):
            log(f"[case] speculative stack grid boxes q={q} preset {v}")
            _twin_equal(gpu, orc, r.scene(), hb, q, rays, v, what="boxes")
    ri = gpu.RayIntersector(gpu.FORMAT_STACK)
    ri.add_object(0, scenes.boxes_scene())
    ri.add_object(3, scenes.icosphere_cloud(count=20, subdivisions=2, seed=3, box=16.0))
    rng = np.random.default_rng(8)
    for k in range(6):
        ang = rng.uniform(0, 2 * np.pi)
        m = np.eye(4)
        m[:3, :3] = [[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]]
        m[:3, 3] = rng.uniform(-10, 10, 3)
        ri.push_entity(k % 2 * 3, m, translucency=0.5 if k == 4 else 0.0)
    ri.buffer_entities()
    ri.buffer_data()
    hi = ri.host_buffers()
    irays = random_rays(20000, [-25, -5, -25], [25, 20, 25], seed=103, toward=[0, 2, 0])
    for flags in (0, gpu.FLAG_IGNORE_TRANSPARENT, gpu.FLAG_TRI0_IS_MISS):
        for q in (0, 1):
            for v in (0, 3):
                log(f"[case] speculative stack grid instances flags={flags} q={q} preset {v}")
                _twin_equal(gpu, orc, ri.scene(), hi, q, irays, v, flags=flags, what="instances")
    ref = orc.trace(1, hi["nodes"], hi["triangles"], hi["vertices"], hi["entities"], 0, irays)
    for n in (0, 1, 63, 65, 257, 1000):
        sub = irays[:max(n, 1)].copy()
        for kernel in (gpu.KERNEL_SPECULATIVE, gpu.KERNEL_SPECULATIVE_REFILL):
            log(f"[case] speculative stack ragged n={n} kernel {kernel}")
            g, _ = _trace(gpu, ri.scene(), 0, sub, kernel, 0, n=n)
            if n == 0:
                assert g.view(np.uint8)[:32].tolist() == [0xAB] * 32
                continue
            assert np.array_equal(g["t"][:n], ref["t"][:n]), f"n={n} kernel={kernel}"
    log("[parity] speculative stack kernels: boxes, instances, flags and ragged batches")


@pytest.mark.parametrize("kind", ["cycle", "bad_child", "overflow"])
def test_guards_equal_twin(gpu, orc, kind):
    v, tris = single_triangle_scene()
    nodes = crafted_stack_chain(70) if kind == "overflow" else crafted_stack(kind)
    ent = orc.identity_entity(nodes.size)
    rays = random_rays(300, [-0.5, -0.5, -1.0], [1.0, 1.0, -0.5], seed=4, toward=[0.3, 0.3, 0.0])
    sc, keep = device_scene(gpu, gpu.FORMAT_STACK, nodes, tris, v, ent)
    for q in (0, 1):
        for var in orc.SBVH2_PRESETS:
            log(f"[case] speculative stack guard {kind} q={q} preset {var}")
            _twin_equal(gpu, orc, sc, {"nodes": nodes, "triangles": tris, "vertices": v, "entities": ent}, q,
                        rays, var, what=kind)
    log(f"[parity] speculative stack guard {kind}: bit-identical to the twin")


def test_errors(gpu, cloud):
    r, _ = cloud
    rays = gpu.DeviceArray.from_host(random_rays(64, [0, 1, 0], [1, 2, 1], seed=1))
    hits = gpu.DeviceArray(gpu.HIT_DTYPE, 64)
    for kernel in (gpu.KERNEL_SPECULATIVE, gpu.KERNEL_SPECULATIVE_REFILL):
        with pytest.raises(gpu.CrtError, match="stack and stackless kernels take variant 0..3"):
            gpu.trace(r.scene(), 0, rays, hits, kernel=kernel, variant=4)
