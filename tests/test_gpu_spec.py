"""GPU parity of the speculative wide walk (crt_trace kernels 4 and 5).

Grid launch (kernel 4): 64 consecutive rays share a wave, so the schedule is
fixed and every record -- t, barycentrics, triangle, mesh, entity, visits,
status and the triangle-test count -- must equal the host twin that replays it
(oracle_trace_wide_spec), for every preset, both widths and both queries.
Refilling launch (kernel 5): which rays share a wave varies from run to run,
but a ray's record does not depend on it (tests/test_oracle_spec_regroup.py),
so every record must equal the same twin run with the refilling preset's
schedule (oracle.SPEC_P_PRESETS), for several grid sizes; the nearest t,
hit / miss and any-hit occlusion also equal the sequential wide walk's
(oracle_trace_wide), the triangle up to exactly equal t.
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


@pytest.fixture(scope="module")
def layouts(gpu):
    from candela_amd import scenes
    r = gpu.RayIntersector(gpu.FORMAT_STACKLESS)
    r.add_object(0, scenes.icosphere_cloud(count=300, subdivisions=3, seed=12, box=40.0))
    r.push_entity(0)
    r.buffer_entities()
    r.buffer_data()
    hb = r.host_buffers()
    return {w: gpu.upload_wide(gpu.FORMAT_STACKLESS, hb, w, 2) for w in (4, 8)}


def _trace(gpu, scene, q, rays, kernel, variant, grid=0, flags=0, n=None):
    dr = gpu.DeviceArray.from_host(rays)
    dh = gpu.DeviceArray(gpu.HIT_DTYPE, max(1, rays.size))
    tt = gpu.DeviceArray(np.uint32, max(1, rays.size))
    dh.fill_bytes(0xAB)
    gpu.trace(scene, q, dr, dh, n=n, kernel=kernel, variant=variant, grid=grid, flags=flags, tri_tests=tt)
    gpu.lib().crt_device_sync()
    return dh.download(), tt.download()


RAYS = None


def _rays(q):
    r = np.concatenate([random_rays(50000, [-5, -5, -5], [45, 45, 45], seed=101, tmax=3.0 if q else 1e6),
                        axis_aligned_rays(3000, [0, 0, 0], [40, 40, 40], seed=102)])
    return r


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("q", [0, 1])
def test_grid_equals_twin(gpu, orc, layouts, width, q):
    dw = layouts[width]
    w = dw.host
    rays = _rays(q)
    for v in range(6):
        log(f"[case] speculative grid width {width} q={q} preset {v}")
        g, gt = _trace(gpu, dw.scene, q, rays, gpu.KERNEL_SPECULATIVE, v)
        o, ot = orc.trace_wide_spec(width, w.groups, w.records, w.entities, q, rays, variant=v, with_tests=True)
        compare(g, o, f"spec grid w{width} q={q} v={v}")
        assert g.tobytes() == o.tobytes()
        assert np.array_equal(gt, ot), f"triangle tests differ (w{width} q={q} v={v})"
    log(f"[parity] speculative grid width {width} q={q}: every preset bit-identical to the twin")


@pytest.mark.parametrize("width", [4, 8])
def test_refill_equals_twin(gpu, orc, layouts, width):
    dw = layouts[width]
    w = dw.host
    for q in (0, 1):
        rays = _rays(q)
        seq = orc.trace_wide(width, w.groups, w.records, w.entities, q, rays)
        twin = {}
        for v in range(6):
            for grid in ((0, 1, 7) if v == 4 else (0,)):
                log(f"[case] speculative refill width {width} q={q} preset {v} grid {grid}")
                g, gt = _trace(gpu, dw.scene, q, rays, gpu.KERNEL_SPECULATIVE_REFILL, v, grid=grid)
                if grid == 0:
                    twin[v] = orc.trace_wide_spec(width, w.groups, w.records, w.entities, q, rays,
                                                  schedule=orc.SPEC_P_PRESETS[v], with_tests=True)
                o, ot = twin[v]
                compare(g, o, f"spec refill w{width} q={q} v={v} grid={grid}")
                assert g.tobytes() == o.tobytes() and np.array_equal(gt, ot), "not bit-identical to the twin"
                assert np.array_equal(g["t"] > 0, seq["t"] > 0), "hit / miss differs"
                assert (g["status"] & 0xFF000000 == 0).all()
                if q == 0:
                    assert np.array_equal(g["t"], seq["t"]), "nearest t differs"
                    same = g["tri"] == seq["tri"]
                    assert (same | (g["t"] == seq["t"])).all() and same.mean() > 0.999
                    for k in ("u", "v", "w"):
                        assert np.array_equal(g[k][same], seq[k][same])
                else:
                    # each accepted occluder lies before its own ray's tmax (3.0 for the random rays,
                    # 1e6 for the axis-aligned ones)
                    hit = g["t"] > 0
                    assert (g["t"][hit] < rays["tmax"][hit]).all()
    log(f"[parity] speculative refill width {width}: every record bit-identical to the twin; nearest t, hits and "
        "occlusion equal the sequential walk")


def test_instances_flags_and_ragged(gpu, orc):
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
    dw = gpu.upload_wide(gpu.FORMAT_STACKLESS, hb, 4, 2)
    w = dw.host
    rays = random_rays(20000, [-25, -5, -25], [25, 20, 25], seed=103, toward=[0, 2, 0])
    for flags in (0, gpu.FLAG_IGNORE_TRANSPARENT, gpu.FLAG_TRI0_IS_MISS):
        for q in (0, 1):
            g, gt = _trace(gpu, dw.scene, q, rays, gpu.KERNEL_SPECULATIVE, 1, flags=flags)
            o, ot = orc.trace_wide_spec(4, w.groups, w.records, w.entities, q, rays, variant=1, flags=flags,
                                        with_tests=True)
            compare(g, o, f"spec instances flags={flags} q={q}")
            assert np.array_equal(gt, ot)
    for n in (0, 1, 63, 65, 257, 1000):
        sub = rays[:max(n, 1)].copy()
        for kernel in (gpu.KERNEL_SPECULATIVE, gpu.KERNEL_SPECULATIVE_REFILL):
            g, _ = _trace(gpu, dw.scene, 0, sub, kernel, 0, n=n)
            if n == 0:
                assert g.view(np.uint8)[:32].tolist() == [0xAB] * 32
                continue
            seq = orc.trace_wide(4, w.groups, w.records, w.entities, 0, sub[:n])
            assert np.array_equal(g["t"][:n], seq["t"]), f"n={n} kernel={kernel}"
    log("[parity] speculative kernels: instances, flags and ragged batches")


def test_errors(gpu, layouts):
    rays = gpu.DeviceArray.from_host(random_rays(64, [0, 1, 0], [1, 2, 1], seed=1))
    hits = gpu.DeviceArray(gpu.HIT_DTYPE, 64)
    with pytest.raises(gpu.CrtError, match="variant 0..5"):
        gpu.trace(layouts[4].scene, 0, rays, hits, kernel=gpu.KERNEL_SPECULATIVE, variant=6)
    s = gpu.Stream()
    with pytest.raises(gpu.CrtError, match="queue"):
        with gpu.Graph(s):
            gpu.trace(layouts[4].scene, 0, rays, hits, kernel=gpu.KERNEL_SPECULATIVE_REFILL, stream=s)
    s.close()
