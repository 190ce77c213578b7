"""Work-queue counters of the persistent kernels never alias.

A persistent launch takes its counters from crt_trace_config.queue or leases
a block from the library's per-device pool; a leased block is reused only
after the launch that held it has completed.  Under stream capture the caller
must bring a block (the graph owns it).  Here a graph-captured refilling
trace is replayed while more than 64 eager refilling traces (the old pool
size) run on a second stream at the same priority; every record of both
streams must equal the host twin of the wide walk.
"""
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
def wide_scene(gpu):
    from candela_amd import scenes
    r = gpu.RayIntersector(gpu.FORMAT_STACKLESS)
    r.add_object(0, scenes.icosphere_cloud(count=120, subdivisions=3, seed=11, box=40.0))
    r.push_entity(0)
    r.buffer_entities()
    r.buffer_data()
    return r, gpu.upload_wide(gpu.FORMAT_STACKLESS, r.host_buffers(), 4, 2)


def test_capture_needs_a_queue(gpu, wide_scene):
    _, dw = wide_scene
    rays = random_rays(4096, [-2, -2, -2], [42, 42, 42], seed=51)
    d_rays = gpu.DeviceArray.from_host(rays)
    hits = gpu.DeviceArray(gpu.HIT_DTYPE, rays.size)
    s = gpu.Stream()
    log("[case] persistent trace captured without a queue block")
    with pytest.raises(gpu.CrtError, match="queue"):
        with gpu.Graph(s):
            gpu.trace(dw.scene, gpu.QUERY_CLOSEST, d_rays, hits, kernel=gpu.KERNEL_PERSISTENT, variant=8, stream=s)
    with pytest.raises(gpu.CrtError, match="queue"):
        with gpu.Graph(s):
            gpu.trace(dw.scene, gpu.QUERY_CLOSEST, d_rays, hits, kernel=gpu.KERNEL_PERSISTENT, variant=0, stream=s)
    with pytest.raises(ValueError):
        gpu.trace(dw.scene, gpu.QUERY_CLOSEST, d_rays, hits, kernel=gpu.KERNEL_PERSISTENT, variant=8, stream=s,
                  queue=gpu.DeviceArray(np.uint8, 64))
    s.sync()
    s.close()


def test_graph_and_eager_persistent_traces_do_not_share_counters(gpu, orc, wide_scene):
    _, dw = wide_scene
    w = dw.host
    rays = random_rays(30000, [-2, -2, -2], [42, 42, 42], seed=52)
    ref_c = orc.trace_wide(w.width, w.groups, w.records, w.entities, 0, rays)
    ref_a = orc.trace_wide(w.width, w.groups, w.records, w.entities, 1, rays)
    d_rays = gpu.DeviceArray.from_host(rays)
    sa, sb = gpu.Stream(), gpu.Stream()
    h_graph = gpu.DeviceArray(gpu.HIT_DTYPE, rays.size)
    h_eager = [gpu.DeviceArray(gpu.HIT_DTYPE, rays.size) for _ in range(2)]
    q = gpu.queue_block()
    log("[case] graph-captured refilling trace with its own queue block")
    gpu.trace(dw.scene, gpu.QUERY_CLOSEST, d_rays, h_graph, kernel=gpu.KERNEL_PERSISTENT, variant=8, stream=sa,
              queue=q)
    sa.sync()
    g = gpu.Graph(sa)
    with g:
        gpu.trace(dw.scene, gpu.QUERY_CLOSEST, d_rays, h_graph, kernel=gpu.KERNEL_PERSISTENT, variant=8, stream=sa,
                  queue=q)
    for rnd in range(3):
        h_graph.fill_bytes(0xFF, stream=sa)
        for h in h_eager:
            h.fill_bytes(0xFF, stream=sb)
        g.launch()
        # more pool leases than the old round-robin pool had blocks, interleaved with graph replays
        for k in range(80):
            query = k % 2
            gpu.trace(dw.scene, query, d_rays, h_eager[query], kernel=gpu.KERNEL_PERSISTENT,
                      variant=8 if k % 4 < 2 else 0, stream=sb)
            if k % 20 == 10:
                g.launch()
        g.launch()
        sa.sync()
        sb.sync()
        got = h_graph.download()
        assert got.tobytes() == ref_c.tobytes(), f"round {rnd}: graph replay records differ from the host twin"
        assert h_eager[0].download().tobytes() == ref_c.tobytes(), f"round {rnd}: eager closest hits differ"
        assert h_eager[1].download().tobytes() == ref_a.tobytes(), f"round {rnd}: eager any hits differ"
        log(f"[parity] round {rnd}: graph replays and 80 eager refilling traces bit-identical to the host twin")
    g.close()
    sa.close()
    sb.close()


def test_paths_under_capture_need_a_queue(gpu, wide_scene):
    from candela_amd import scenes
    _, dw = wide_scene
    iv, ip = scenes.camera_inverses(eye=(-5, 20, -5), center=(20, 20, 20), fovy_deg=60.0, aspect=1.0)
    p = gpu.path_params(32, 32, 1, iv, ip, seed=3)
    out = gpu.DeviceArray(gpu.PATH_SAMPLE_DTYPE, 32 * 32)
    s = gpu.Stream()
    gpu.trace_paths(dw.scene, p, out, stream=s)
    s.sync()
    eager = out.download()
    with pytest.raises(gpu.CrtError, match="queue"):
        with gpu.Graph(s):
            gpu.trace_paths(dw.scene, p, out, stream=s)
    q = gpu.queue_block()
    g = gpu.Graph(s)
    with g:
        gpu.trace_paths(dw.scene, p, out, stream=s, queue=q)
    for _ in range(3):
        out.fill_bytes(0xFF, stream=s)
        g.launch()
        s.sync()
        assert out.download().tobytes() == eager.tobytes()
    g.close()
    s.close()
