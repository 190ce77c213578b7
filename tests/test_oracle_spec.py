"""The host twin of the speculative wide walk (oracle_trace_wide_spec), without a GPU.

Whatever the wave schedule, culling only ever uses a tmax at least as large
as the sequential walk's, so a closest-hit walk must find the nearest t the
sequential wide walk (oracle_trace_wide) finds, on every ray; the triangle may
differ only where both walks report the same t (an exact tie between
triangles), and then it is one the sequential walk can reach; any-hit walks
must agree on occlusion.  Checked over every preset schedule and over the
smallest queues (a lane that queued any leaf waits for the leaf phase), on
clouds with random rays, on instanced entities with IgnoreTransparent, and
against brute force.
"""
import numpy as np
import pytest

from oracle_cases import axis_aligned_rays, random_rays

SCHEDULES = [None, (8, 4, 4), (8, 1, 4), (8, 16, 16)]


def _cloud(crt, orc, width):
    import oracle
    from candela_amd import scenes
    b = crt.build_bvh(0, scenes.icosphere_cloud(count=120, subdivisions=2, seed=13, box=30.0))
    ent = oracle.identity_entity(b.nodes.size)
    return b, crt.wide_build(0, b.nodes, b.triangles, b.vertices, ent, width, 2)


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("schedule", SCHEDULES)
def test_spec_nearest_t_and_occlusion(crt, orc, width, schedule):
    b, w = _cloud(crt, orc, width)
    rays = np.concatenate([random_rays(8000, [-3, -3, -3], [33, 33, 33], seed=91),
                           axis_aligned_rays(1000, [0, 0, 0], [30, 30, 30], seed=92)])
    variants = range(6) if schedule is None else [0]
    for v in variants:
        seq = orc.trace_wide(width, w.groups, w.records, w.entities, 0, rays)
        spec, tests = orc.trace_wide_spec(width, w.groups, w.records, w.entities, 0, rays, variant=v,
                                          schedule=schedule, with_tests=True)
        assert np.array_equal(spec["t"], seq["t"]), f"nearest t differs (variant {v}, schedule {schedule})"
        same = spec["tri"] == seq["tri"]
        assert (same | (spec["t"] == seq["t"])).all()
        assert same.mean() > 0.999
        for k in ("u", "v", "w"):
            assert np.array_equal(spec[k][same], seq[k][same])
        assert (spec["status"] & 0xFF000000 == 0).all()
        hit = spec["t"] > 0
        assert hit.mean() > 0.2 and (tests > 0).mean() > 0.2
        rays_any = rays.copy()
        rays_any["tmax"] = np.float32(2.5)
        sa = orc.trace_wide(width, w.groups, w.records, w.entities, 1, rays_any)
        pa = orc.trace_wide_spec(width, w.groups, w.records, w.entities, 1, rays_any, variant=v, schedule=schedule)
        assert np.array_equal(pa["t"] > 0, sa["t"] > 0), "occlusion differs"
        assert (pa["t"][pa["t"] > 0] < 2.5).all()


def test_spec_against_brute_force(crt, orc):
    import oracle
    from candela_amd import scenes
    b = crt.build_bvh(0, scenes.boxes_scene())
    ent = oracle.identity_entity(b.nodes.size)
    w = crt.wide_build(0, b.nodes, b.triangles, b.vertices, ent, 4, 2)
    rays = random_rays(6000, [-4, 0.2, -4], [20, 8, 20], seed=93)
    spec = orc.trace_wide_spec(4, w.groups, w.records, w.entities, 0, rays, variant=1)
    bt, _ = orc.bruteforce_closest(b.triangles, b.vertices, ent, rays)
    assert np.array_equal(spec["t"], bt)


def test_spec_instances_and_flags(crt, orc):
    import oracle
    from candela_amd import scenes
    r = crt.RayIntersector(0)
    r.add_object(0, scenes.boxes_scene())
    r.add_object(3, scenes.icosphere_cloud(count=20, subdivisions=2, seed=3, box=16.0))
    rng = np.random.default_rng(5)
    for k in range(6):
        ang = rng.uniform(0, 2 * np.pi)
        m = np.eye(4)
        m[:3, :3] = [[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]]
        m[:3, 3] = rng.uniform(-10, 10, 3)
        r.push_entity(k % 2 * 3, m, translucency=0.5 if k == 4 else 0.0)
    hb = r.host_buffers() if False else None
    import ctypes
    # the pending entities are the context's host copy once buffered: build the record here instead
    ents = np.zeros(6, crt.ENTITY_DTYPE)
    b0 = crt.build_bvh(0, scenes.boxes_scene())
    b3 = crt.build_bvh(0, scenes.icosphere_cloud(count=20, subdivisions=2, seed=3, box=16.0), t_offset=b0.triangles.size)
    nodes = np.concatenate([b0.nodes, b3.nodes])
    v3 = b3.triangles.copy()
    v3["packed"][:, :3] += b0.vertices.size
    tris = np.concatenate([b0.triangles, v3])
    verts = np.concatenate([b0.vertices, b3.vertices])
    rng = np.random.default_rng(5)
    for k in range(6):
        ang = rng.uniform(0, 2 * np.pi)
        m = np.eye(4)
        m[:3, :3] = [[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]]
        m[:3, 3] = rng.uniform(-10, 10, 3)
        off, cnt = (0, b0.nodes.size) if k % 2 == 0 else (b0.nodes.size, b3.nodes.size)
        mm = np.ascontiguousarray(m.astype(np.float32).T.reshape(16))
        crt.check(crt.lib().crt_make_entity(mm.ctypes.data, 0.0, 0.5 if k == 4 else 0.0, off, cnt,
                                            ents[k:k + 1].ctypes.data))
    w = crt.wide_build(0, nodes, tris, verts, ents, 4, 2)
    rays = random_rays(8000, [-25, -5, -25], [25, 20, 25], seed=94, toward=[0, 2, 0])
    for flags in (0, crt.FLAG_IGNORE_TRANSPARENT, crt.FLAG_TRI0_IS_MISS):
        seq = orc.trace_wide(4, w.groups, w.records, w.entities, 0, rays, flags=flags)
        spec = orc.trace_wide_spec(4, w.groups, w.records, w.entities, 0, rays, flags=flags, variant=0)
        assert np.array_equal(spec["t"], seq["t"])
        assert np.array_equal(spec["entity"][spec["t"] > 0], seq["entity"][seq["t"] > 0])
        sa = orc.trace_wide(4, w.groups, w.records, w.entities, 1, rays, flags=flags)
        pa = orc.trace_wide_spec(4, w.groups, w.records, w.entities, 1, rays, flags=flags, variant=0)
        assert np.array_equal(pa["t"] > 0, sa["t"] > 0)
