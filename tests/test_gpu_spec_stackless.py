"""GPU parity of the speculative walk of the reference's stackless layout (crt_trace kernels 4 and 5 over
CRT_FORMAT_STACKLESS).

Grid launch (kernel 4): 64 consecutive rays share a wave, so the schedule is fixed and every record -- t,
barycentrics, triangle, mesh, entity, visits, status and the triangle-test count -- must equal the host twin that
replays it (oracle_trace_stackless_spec), for every preset and both queries, on a cloud, on instanced entities
with the scene flags, and on the boxes scene with touching boxes.
Refilling launch (kernel 5): which rays share a wave varies from run to run, but a ray's record does not depend on
it (tests/test_oracle_spec_regroup.py), so every record must equal the twin run with the refilling preset's schedule
(oracle.SBVH2_P_PRESETS); on the cloud (no touching boxes) the nearest t, hit / miss and any-hit occlusion also equal
the reference walk's (oracle_trace) and the triangle may differ only at exactly equal t.
"""
import sys

import numpy as np
import pytest

from oracle_cases import axis_aligned_rays, random_rays
from test_gpu_trace import compare

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
    r = gpu.RayIntersector(gpu.FORMAT_STACKLESS)
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


@pytest.mark.parametrize("q", [0, 1])
def test_grid_equals_twin(gpu, orc, cloud, q):
    r, hb = cloud
    rays = _rays(q)
    for v in orc.SBVH2_PRESETS:
        log(f"[case] speculative stackless grid q={q} preset {v}")
        g, gt = _trace(gpu, r.scene(), q, rays, gpu.KERNEL_SPECULATIVE, v)
        o, ot = orc.trace_stackless_spec(hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], q, rays,
                                         variant=v, with_tests=True)
        compare(g, o, f"spec stackless grid q={q} v={v}")
        assert g.tobytes() == o.tobytes()
        assert np.array_equal(gt, ot), f"triangle tests differ (q={q} v={v})"
    log(f"[parity] speculative stackless grid q={q}: every preset bit-identical to the twin")


def test_refill_nearest_and_occlusion(gpu, orc, cloud):
    r, hb = cloud
    for q in (0, 1):
        rays = _rays(q)
        ref = orc.trace(0, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], q, rays)
        for v in orc.SBVH2_PRESETS:
            for grid in ((0, 1, 7) if v == 1 else (0,)):
                log(f"[case] speculative stackless refill q={q} preset {v} grid {grid}")
                g, gt = _trace(gpu, r.scene(), q, rays, gpu.KERNEL_SPECULATIVE_REFILL, v, grid=grid)
                o, ot = orc.trace_stackless_spec(hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], q, rays,
                                 schedule=orc.SBVH2_P_PRESETS[v], with_tests=True)
                compare(g, o, f"refill q={q} v={v} grid={grid}")
                assert g.tobytes() == o.tobytes() and np.array_equal(gt, ot), "not bit-identical to the twin"
                assert np.array_equal(g["t"] > 0, ref["t"] > 0), "hit / miss differs"
                assert (g["status"] & 0xFF000000 == 0).all()
                if q == 0:
                    assert np.array_equal(g["t"], ref["t"]), "nearest t differs"
                    same = g["tri"] == ref["tri"]
                    assert (same | (g["t"] == ref["t"])).all() and same.mean() > 0.999
                    for k in ("u", "v", "w"):
                        assert np.array_equal(g[k][same], ref[k][same])
                else:
                    hit = g["t"] > 0
                    assert (g["t"][hit] < rays["tmax"][hit]).all()
    log("[parity] speculative stackless refill: every record bit-identical to the twin; nearest t, hits and occlusion "
        "equal the reference walk")


def test_instances_flags_touching_boxes_and_ragged(gpu, orc):
    from candela_amd import scenes
    r = gpu.RayIntersector(gpu.FORMAT_STACKLESS)
    r.add_object(0, scenes.boxes_scene())
    r.add_object(3, scenes.icosphere_cloud(count=20, subdivisions=2, seed=3, box=16.0))
    rng = np.random.default_rng(8)
    for k in range(6):
        ang = rng.uniform(0, 2 * np.pi)
        m = np.eye(4)
        m[:3, :3] = [[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]]
        m[:3, 3] = rng.uniform(-10, 10, 3)
        r.push_entity(k % 2 * 3, m, translucency=0.5 if k == 4 else 0.0)
    r.buffer_entities()
    r.buffer_data()
    hb = r.host_buffers()
    rays = random_rays(20000, [-25, -5, -25], [25, 20, 25], seed=103, toward=[0, 2, 0])
    for flags in (0, gpu.FLAG_IGNORE_TRANSPARENT, gpu.FLAG_TRI0_IS_MISS):
        for q in (0, 1):
            for v in (0, 3):
                g, gt = _trace(gpu, r.scene(), q, rays, gpu.KERNEL_SPECULATIVE, v, flags=flags)
                o, ot = orc.trace_stackless_spec(hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], q,
                                                 rays, variant=v, flags=flags, with_tests=True)
                compare(g, o, f"spec stackless instances flags={flags} q={q} v={v}")
                assert np.array_equal(gt, ot)
    ref = orc.trace(0, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], 0, rays)
    for n in (0, 1, 63, 65, 257, 1000):
        sub = rays[:max(n, 1)].copy()
        for kernel in (gpu.KERNEL_SPECULATIVE, gpu.KERNEL_SPECULATIVE_REFILL):
            g, _ = _trace(gpu, r.scene(), 0, sub, kernel, 0, n=n)
            if n == 0:
                assert g.view(np.uint8)[:32].tolist() == [0xAB] * 32
                continue
            # touching boxes: a hit may only be nearer than the reference walk's
            near = lambda t: np.where(t > 0, t, np.inf)
            assert (near(g["t"][:n]) <= near(ref["t"][:n])).all(), f"n={n} kernel={kernel}"
    log("[parity] speculative stackless kernels: instances, flags, touching boxes and ragged batches")


def test_errors(gpu, cloud):
    r, _ = cloud
    rays = gpu.DeviceArray.from_host(random_rays(64, [0, 1, 0], [1, 2, 1], seed=1))
    hits = gpu.DeviceArray(gpu.HIT_DTYPE, 64)
    for kernel in (gpu.KERNEL_SPECULATIVE, gpu.KERNEL_SPECULATIVE_REFILL):
        with pytest.raises(gpu.CrtError, match="stackless kernels take variant 0..3"):
            gpu.trace(r.scene(), 0, rays, hits, kernel=kernel, variant=4)
