"""GPU parity: the HIP kernels of libcandela_rt against the host oracle on identical inputs.

Every comparison runs the same BVH buffers and the same ray records through
crt_trace / crt_intersect / crt_intersect_primary (GPU) and oracle_trace (CPU).
Both do the same FP32 operations in the same order, so the bar is exact
equality of t, u, v, w, triangle, mesh, entity and status on every ray; the
path's stated tolerance (>= 99.999 % hit-ID agreement, t and barycentrics within
1e-5 relative) is asserted as well so that a failure reports both.
"""
import sys

import numpy as np
import pytest

from oracle_cases import (axis_aligned_rays, crafted_stack_chain, crafted_stackless, random_rays,
                          single_triangle_scene)

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu(crt):
    if crt.device_count() < 1:
        pytest.fail("no HIP device visible")
    return crt


def compare(gh, oh, label=""):
    """Asserts exact agreement and the tolerance bar; returns the exact-match fraction."""
    assert gh.shape == oh.shape
    n = gh.size
    ids = (gh["tri"] == oh["tri"]) & (gh["mesh"] == oh["mesh"]) & (gh["entity"] == oh["entity"])
    id_rate = ids.mean() if n else 1.0
    both = ids & (oh["t"] > 0)
    rel = lambda a, b: np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
    assert id_rate >= 0.99999, f"{label}: hit-ID agreement {id_rate}"
    for k in ("t", "u", "v", "w"):
        assert (rel(gh[k][both], oh[k][both]) <= 1e-5).all(), f"{label}: {k} beyond 1e-5 relative"
    exact = np.ones(n, bool)
    for k in ("t", "u", "v", "w"):
        exact &= (gh[k] == oh[k]) | (np.isnan(gh[k]) & np.isnan(oh[k]))
    exact &= ids & (gh["status"] == oh["status"])
    rate = exact.mean() if n else 1.0
    log(f"[parity] {label}: n={n} id={id_rate:.7f} exact={rate:.7f} hit={np.mean(oh['t'] > 0):.3f}")
    assert rate == 1.0, f"{label}: {np.count_nonzero(~exact)} rays differ bitwise"
    return rate


def device_scene(crt, fmt, nodes, tris, verts, ents):
    d = {k: crt.DeviceArray.from_host(v) for k, v in
         (("nodes", nodes), ("tris", tris), ("verts", verts), ("ents", ents))}
    s = crt.SceneView(fmt, ents.size, nodes.size, d["nodes"].ptr, d["tris"].ptr, d["verts"].ptr, d["ents"].ptr)
    return s, d


def gpu_trace(crt, scene, query, rays, flags=0, block=0, lds=0, kernel=0, variant=0, grid=0):
    dr = crt.DeviceArray.from_host(rays)
    dh = crt.DeviceArray(crt.HIT_DTYPE, rays.size)
    crt.trace(scene, query, dr, dh, flags=flags, block_size=block, lds_stack=lds, kernel=kernel, variant=variant,
              grid=grid)
    crt.lib().crt_device_sync()
    return dh.download()


# (format, kernel) pairs: the one-ray-per-lane, persistent and wave kernels
# over both layouts of the reference.
WALKERS = [(0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3)]


@pytest.mark.parametrize("fmt", [0, 1])
def test_primary_boxes_512(gpu, orc, fmt):
    """Config 1: IntersectPrimary over 512x512 pixels of the 16x16 boxes scene."""
    from candela_amd import scenes
    log(f"[case] primary boxes 512 fmt={fmt}")
    r = gpu.RayIntersector(fmt)
    r.add_object(0, scenes.boxes_scene())
    r.push_entity(0)
    r.buffer_entities()
    r.buffer_data()
    iv, ip = scenes.camera_inverses(aspect=1.0, **scenes.BOXES_CAMERA)
    drays, dhits = r.intersect_primary(512, 512, iv, ip)
    gpu.lib().crt_device_sync()
    grays, ghits = drays.download(), dhits.download()
    hb = r.host_buffers()
    orays = orc.primary_rays(512, 512, iv.T.reshape(16), ip.T.reshape(16))
    assert grays.tobytes() == orays.tobytes(), "primary ray generation differs"
    oh = orc.trace(fmt, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], 0, orays)
    compare(ghits, oh, f"primary fmt={fmt}")
    assert 0.5 < np.mean(oh["t"] > 0) < 1.0


@pytest.mark.parametrize("fmt,kernel", WALKERS)
@pytest.mark.parametrize("query", [0, 1])
def test_random_rays_cloud(gpu, orc, fmt, kernel, query):
    from candela_amd import scenes
    log(f"[case] random rays cloud fmt={fmt} kernel={kernel} query={query}")
    b = gpu.build_bvh(fmt, scenes.icosphere_cloud(count=300, subdivisions=3, seed=12, box=40.0))
    import oracle
    ents = oracle.identity_entity(b.nodes.size)
    s, keep = device_scene(gpu, fmt, b.nodes, b.triangles, b.vertices, ents)
    rays = np.concatenate([random_rays(60000, [-5, -5, -5], [45, 45, 45], seed=21, tmax=3.0 if query else 1e6),
                           axis_aligned_rays(4000, [0, 0, 0], [40, 40, 40], seed=22)])
    gh = gpu_trace(gpu, s, query, rays, kernel=kernel)
    oh = orc.trace(fmt, b.nodes, b.triangles, b.vertices, ents, query, rays)
    compare(gh, oh, f"cloud fmt={fmt} kernel={kernel} q={query}")


@pytest.mark.parametrize("fmt", [0, 1])
def test_instances_and_flags(gpu, orc, fmt):
    """Several transformed entities of two objects, IgnoreTransparent and the tri-0 quirk."""
    from candela_amd import scenes
    log(f"[case] instances fmt={fmt}")
    r = gpu.RayIntersector(fmt)
    r.add_object(0, scenes.boxes_scene())
    r.add_object(3, scenes.icosphere_cloud(count=20, subdivisions=2, seed=3, box=16.0))
    rng = np.random.default_rng(5)
    for k in range(6):
        ang = rng.uniform(0, 2 * np.pi)
        m = np.eye(4)
        m[:3, :3] = [[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]]
        m[:3, :3] *= rng.uniform(0.5, 1.5)
        m[:3, 3] = rng.uniform(-10, 10, 3)
        r.push_entity(k % 2 * 3, m, emissive=0.1 * k, translucency=0.5 if k == 4 else 0.0)
    r.buffer_entities()
    r.buffer_data()
    hb = r.host_buffers()
    rays = random_rays(30000, [-25, -5, -25], [25, 20, 25], seed=23, toward=[0, 2, 0])
    dr = gpu.DeviceArray.from_host(rays)
    dh = gpu.DeviceArray(gpu.HIT_DTYPE, rays.size)
    for flags in (0, gpu.FLAG_IGNORE_TRANSPARENT, gpu.FLAG_TRI0_IS_MISS, 3):
        for query in (0, 1):
            r.intersect(query, dr, dh, flags=flags)
            gpu.lib().crt_device_sync()
            oh = orc.trace(fmt, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], query, rays,
                           flags=flags)
            compare(dh.download(), oh, f"instances fmt={fmt} flags={flags} q={query}")
            gpu.trace(r.scene(), query, dr, dh, flags=flags, kernel=gpu.KERNEL_THREAD_PER_RAY)
            gpu.lib().crt_device_sync()
            compare(dh.download(), oh, f"instances thread-per-ray fmt={fmt} flags={flags} q={query}")


def test_launch_shapes_do_not_change_results(gpu, orc):
    from candela_amd import scenes
    log("[case] launch shapes")
    import oracle
    for fmt in (0, 1):
        b = gpu.build_bvh(fmt, scenes.icosphere_cloud(count=100, subdivisions=3, seed=31, box=30.0))
        ents = oracle.identity_entity(b.nodes.size)
        s, keep = device_scene(gpu, fmt, b.nodes, b.triangles, b.vertices, ents)
        rays = random_rays(20001, [-2, -2, -2], [32, 32, 32], seed=32)
        oh = orc.trace(fmt, b.nodes, b.triangles, b.vertices, ents, 0, rays)
        shapes = [(64, 8), (128, 16), (256, 32), (512, 16), (1024, 8), (256, 64), (256, 48)] if fmt else \
            [(64, 0), (192, 0), (256, 0), (1024, 0)]
        for block, lds in shapes:
            gh = gpu_trace(gpu, s, 0, rays, block=block, lds=lds, kernel=1)
            compare(gh, oh, f"fmt={fmt} block={block} lds={lds}")
        if fmt == 0:
            for q in (0, 1):
                oq = oh if q == 0 else orc.trace(fmt, b.nodes, b.triangles, b.vertices, ents, 1, rays)
                for variant in range(6):
                    gh = gpu_trace(gpu, s, q, rays, kernel=3, variant=variant)
                    compare(gh, oq, f"wave q={q} variant={variant}")
                    for grid in (0, 1, 7, 33):
                        gh = gpu_trace(gpu, s, q, rays, kernel=2, variant=variant, grid=grid)
                        compare(gh, oq, f"persistent q={q} variant={variant} grid={grid}")


def test_guards_on_gpu(gpu, orc):
    """Crafted buffers: the GPU reports the same status bits as the oracle (all accesses stay in bounds)."""
    log("[case] guards")
    import oracle
    v, tris = single_triangle_scene()
    ray = random_rays(64, [0.1, 0.1, 1], [0.3, 0.3, 1], seed=0)
    ray["dir"] = [0, 0, -1]
    cases = [(0, crafted_stackless("bad_link"), 1 << 26), (0, crafted_stackless("cycle"), 1 << 24),
             (1, crafted_stack_chain(70), 1 << 25)]
    for fmt, nodes, flag in cases:
        ents = oracle.identity_entity(nodes.size)
        s, keep = device_scene(gpu, fmt, nodes, tris, v, ents)
        for kernel, lds in (((1, 0), (2, 0), (3, 0)) if fmt == 0 else ((1, 8), (1, 16), (1, 64), (2, 0), (3, 0))):
            gh = gpu_trace(gpu, s, 0, ray, lds=lds, kernel=kernel)
            oh = orc.trace(fmt, nodes, tris, v, ents, 0, ray)
            assert (oh["status"] & 0xFF000000 == flag).all()
            compare(gh, oh, f"guard fmt={fmt} flag={flag:#x} kernel={kernel} lds={lds}")


def test_empty_and_ragged_batches(gpu, orc):
    from candela_amd import scenes
    log("[case] empty/ragged")
    import oracle
    b = gpu.build_bvh(0, scenes.boxes_scene())
    ents = oracle.identity_entity(b.nodes.size)
    s, keep = device_scene(gpu, 0, b.nodes, b.triangles, b.vertices, ents)
    for n in (0, 1, 63, 65, 257, 1000):
        rays = random_rays(max(n, 1), [-4, 0.2, -4], [20, 8, 20], seed=40 + n)[:n]
        dr = gpu.DeviceArray(gpu.RAY_DTYPE, max(n, 1))
        if n:
            dr.upload(np.concatenate([rays]))
        dh = gpu.DeviceArray(gpu.HIT_DTYPE, max(n, 1))
        for kernel in (1, 2, 3):
            dh.fill_bytes(0x7F)
            gpu.trace(s, 0, dr, dh, n=n, kernel=kernel)
            gpu.lib().crt_device_sync()
            out = dh.download()
            if n == 0:
                assert out.view(np.uint8).tolist() == [0x7F] * 32   # nothing written
                continue
            compare(out[:n], orc.trace(0, b.nodes, b.triangles, b.vertices, ents, 0, rays), f"n={n} kernel={kernel}")
