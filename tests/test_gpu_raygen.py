"""GPU parity of the wavefront ray generators (crt_gen_diffuse_rays,
crt_gen_specular_rays, crt_gen_probe_rays) against their host twin.

The G-buffer is the primary closest hits of a 333 x 211 image (not a multiple
of the 256-record block) over two objects with three entities, one scaled and
rotated; the hits come from the BVH2 view and from the BVH4 view (which gives
positions through its triangle records).  Every generated ray must be
bit-identical to oracle/oracle_raygen.c, in the same compacted order, and the
device count must equal the number of records with t > 0.  Specular rays use a
roughness table with mirror (0 and 0.005), rough and over-one entries and mesh
ids beyond the table.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup(crt):
    from candela_amd import scenes
    if crt.device_count() < 1:
        pytest.fail("no HIP device visible")
    log("[case] raygen scene")
    r = crt.RayIntersector(0)
    r.add_object(0, scenes.boxes_scene())
    r.add_object(1, [(v, i, 2 + m) for v, i, m in scenes.icosphere_cloud(count=30, subdivisions=2, seed=9, box=12.0)])
    c, s = np.cos(0.4), np.sin(0.4)
    rot = np.array([[1.5 * c, 0, s, 2.0], [0, 0.75, 0, 4.0], [-s, 0, c, 1.0], [0, 0, 0, 1]], np.float32)
    r.push_entity(0)
    r.push_entity(1, rot)
    r.push_entity(1)
    r.buffer_entities()
    r.buffer_data()
    r.build_wide(4, 2)
    W, H = 333, 211
    iv, ip = scenes.camera_inverses(aspect=W / H, **scenes.BOXES_CAMERA)
    drays, dhits = r.intersect_primary(W, H, iv, ip)
    crt.lib().crt_device_sync()
    return r, drays, dhits, W * H


def _host(r):
    hb = r.host_buffers()
    return hb["triangles"], hb["vertices"], hb["entities"]


@pytest.mark.parametrize("view", ["bvh2", "wide"])
def test_diffuse_rays(crt, orc, setup, view):
    r, drays, dhits, n = setup
    log(f"[case] diffuse rays view={view}")
    scene = r.scene() if view == "bvh2" else r.wide_scene()
    gen = crt.RayGen(n)
    ao, gi = crt.DeviceArray(crt.RAY_DTYPE, n), crt.DeviceArray(crt.RAY_DTYPE, n)
    ao.fill_bytes(0xCD)
    gi.fill_bytes(0xCD)
    p = crt.raygen_params(seed=5, offset=1e-4)
    gen.diffuse(scene, drays, dhits, p, ao, gi)
    k = gen.last_count()
    rays, hits = drays.download(), dhits.download()
    tris, verts, ents = _host(r)
    oao, ogi = orc.gen_diffuse_rays(p, rays, hits, ents, tris=tris, verts=verts)
    assert k == oao.size == int((hits["t"] > 0).sum()) > 10000
    assert ao.download()[:k].tobytes() == oao.tobytes()
    assert gi.download()[:k].tobytes() == ogi.tobytes()
    log(f"[parity] diffuse {view}: {k} AO + {k} GI rays of {n} records bit-identical")
    # one output only
    gi.fill_bytes(0)
    gen.diffuse(scene, drays, dhits, p, None, gi)
    assert gen.last_count() == k and gi.download()[:k].tobytes() == ogi.tobytes()


def test_specular_rays(crt, orc, setup):
    r, drays, dhits, n = setup
    log("[case] specular rays")
    rng = np.random.default_rng(8)
    table = rng.uniform(0.02, 1.3, size=20).astype(np.float32)   # meshes 20.. are beyond the table
    table[0], table[2], table[3] = 0.0, 0.005, 1.0
    drough = crt.DeviceArray.from_host(table)
    gen = crt.RayGen(n)
    out = crt.DeviceArray(crt.RAY_DTYPE, n)
    rays, hits = drays.download(), dhits.download()
    tris, verts, ents = _host(r)
    for scale, offs in ((0.9, (0.05, 0.1)), (1.0, (1e-4, 2e-4))):
        p = crt.raygen_params(seed=6, spec_offset=offs, roughness_scale=scale, roughness=drough)
        gen.specular(r.scene(), drays, dhits, p, out)
        k = gen.last_count()
        o = orc.gen_specular_rays(p, rays, hits, ents, tris=tris, verts=verts, roughness=table)
        assert k == o.size and out.download()[:k].tobytes() == o.tobytes()
    p = crt.raygen_params(seed=7)                                  # no table: every mesh 1.0
    gen.specular(r.wide_scene(), drays, dhits, p, out)
    k = gen.last_count()
    o = orc.gen_specular_rays(p, rays, hits, ents, tris=tris, verts=verts)
    assert k == o.size and out.download()[:k].tobytes() == o.tobytes()
    log(f"[parity] specular: {k} rays bit-identical (three parameter sets)")


def test_probe_rays(crt, orc):
    log("[case] probe rays")
    for res in ((37, 5, 11), (1, 1, 1), (64, 32, 64)):
        g = crt.probe_grid(res, (10.0, -3.0, 7.5), (32.0, 8.0, 20.0), seed=3, tmax=250.0)
        n = int(np.prod(res))
        out = crt.DeviceArray(crt.RAY_DTYPE, n)
        crt.gen_probe_rays(g, out)
        assert out.download().tobytes() == orc.gen_probe_rays(g).tobytes(), res


def test_edges_and_errors(crt, setup):
    import ctypes
    r, drays, dhits, n = setup
    log("[case] raygen edges")
    gen = crt.RayGen(n)
    ao = crt.DeviceArray(crt.RAY_DTYPE, n)
    p = crt.raygen_params()
    gen.diffuse(r.scene(), drays, dhits, p, ao, None, n=0)
    assert gen.last_count() == 0
    misses = crt.DeviceArray.from_host(_miss_hits(crt, 1000))
    gen.diffuse(r.scene(), drays, misses, p, ao, None, n=1000)
    assert gen.last_count() == 0
    scene = r.scene()

    def call(ws, ws_bytes, count):
        crt.check(crt.lib().crt_gen_diffuse_rays(ctypes.byref(scene), drays.ptr, dhits.ptr, n, crt.ptr(p), ao.ptr,
                                                 None, count, ws, ws_bytes, None))

    small = crt.DeviceArray(np.uint8, 64)
    for args in ((small.ptr, small.nbytes, gen.count.ptr),                            # too small
                 (gen.workspace.ptr + 4, gen.workspace.nbytes - 4, gen.count.ptr),    # misaligned
                 (gen.workspace.ptr, gen.workspace.nbytes, None)):                    # no counter
        with pytest.raises(crt.CrtError):
            call(*args)
    with pytest.raises(crt.CrtError):
        crt.gen_probe_rays(crt.probe_grid((0, 4, 4), (0, 0, 0), (1, 1, 1)), ao)


def _miss_hits(crt, n):
    h = np.zeros(n, crt.HIT_DTYPE)
    h["t"] = h["u"] = h["v"] = h["w"] = -1.0
    h["tri"] = h["mesh"] = h["entity"] = -1
    return h
