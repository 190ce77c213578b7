"""The host twin of the speculative walk of the reference's stack layout (oracle_trace_stack_spec), without a GPU.

The walk is the reference's stack walk (TraverseBVHStack.glsl:168-324, :509-657) with a node's non-empty leaf
children queued instead of tested.  The next entry is decided from the inner children's box distances, which do not
depend on TMax, and boxes are culled with the TMax of the last leaf phase, which is never below the reference
walk's at the same node: the reference walk's sequence of entries only gains whole subtrees.  So the nearest t equals
the reference walk's, the triangle differs only at an exact tie of t, node visits are never fewer, and an any-hit
walk reports the reference walk's first accepted triangle (its queue order is the reference walk's test order up
to that triangle).  The stack layout's RayBounds keeps a box whose entry equals TMax, so where boxes touch no
nearer triangle is skipped: the nearest t also equals brute force on the boxes scene.  Checked over every preset
and small schedules, on clouds with random and axis-parallel rays, on the boxes scene, on instanced entities with
the scene flags and on the crafted guard buffers.
"""
import numpy as np
import pytest

from oracle_cases import (axis_aligned_rays, crafted_stack, crafted_stack_chain, instanced_bvh2, random_rays,
                          single_triangle_scene)

SCHEDULES = [None, (1, 2), (4, 3), (64, 16)]


def _scene(crt, orc, name):
    from candela_amd import scenes
    if name == "cloud":
        b = crt.build_bvh(1, scenes.icosphere_cloud(count=120, subdivisions=2, seed=13, box=30.0))
        rays = np.concatenate([random_rays(8000, [-3, -3, -3], [33, 33, 33], seed=91),
                               axis_aligned_rays(1000, [0, 0, 0], [30, 30, 30], seed=92)])
    else:
        b = crt.build_bvh(1, scenes.boxes_scene())
        rays = random_rays(20000, [-4, 0.2, -4], [20, 8, 20], seed=3)
    return b, orc.identity_entity(b.nodes.size), rays


def _check(orc, b, ent, rays, ref, ref_any, bt, variant, schedule):
    h = orc.trace_stack_spec(b.nodes, b.triangles, b.vertices, ent, 0, rays, variant=variant, schedule=schedule)
    assert np.array_equal(h["t"], ref["t"]), f"nearest t differs (variant {variant}, schedule {schedule})"
    assert np.array_equal(h["t"], bt), "nearest t differs from brute force"
    same = h["tri"] == ref["tri"]
    assert (same | (h["t"] == ref["t"])).all()
    for k in ("u", "v", "w"):
        assert np.array_equal(h[k][same], ref[k][same])
    vis, rvis = h["status"] & 0xFFFFFF, ref["status"] & 0xFFFFFF
    assert (vis >= rvis).all(), "the speculative walk skipped an entry the reference walk visits"
    assert (h["status"] >> 24 == 0).all()
    a = orc.trace_stack_spec(b.nodes, b.triangles, b.vertices, ent, 1, rays, variant=variant, schedule=schedule)
    assert np.array_equal(a["t"], ref_any["t"]), "any hit: not the reference walk's first accepted triangle"
    assert np.array_equal(a["tri"], ref_any["tri"])
    return h


@pytest.mark.parametrize("schedule", SCHEDULES)
@pytest.mark.parametrize("scene", ["cloud", "boxes"])
def test_equals_reference_walk(crt, orc, scene, schedule):
    b, ent, rays = _scene(crt, orc, scene)
    bt, _ = orc.bruteforce_closest(b.triangles, b.vertices, ent, rays)
    ref = orc.trace(1, b.nodes, b.triangles, b.vertices, ent, 0, rays)
    ref_any = orc.trace(1, b.nodes, b.triangles, b.vertices, ent, 1, rays)
    assert (ref["t"] > 0).mean() > 0.1
    for v in (sorted(orc.SBVH2_PRESETS) if schedule is None else [0]):
        _check(orc, b, ent, rays, ref, ref_any, bt, v, schedule)


def test_instances_and_flags(crt, orc):
    nodes, tris, verts, ents = instanced_bvh2(crt, 1)
    rays = random_rays(6000, [-25, -5, -25], [25, 20, 25], seed=103, toward=[0, 2, 0])
    for flags in (0, crt.FLAG_IGNORE_TRANSPARENT, crt.FLAG_TRI0_IS_MISS):
        ref = orc.trace(1, nodes, tris, verts, ents, 0, rays, flags=flags)
        refa = orc.trace(1, nodes, tris, verts, ents, 1, rays, flags=flags)
        for v in (0, 3):
            h = orc.trace_stack_spec(nodes, tris, verts, ents, 0, rays, variant=v, flags=flags)
            assert np.array_equal(h["t"], ref["t"])
            same = h["tri"] == ref["tri"]
            assert (same | (h["t"] == ref["t"])).all()
            assert np.array_equal(h["entity"], ref["entity"])
            a = orc.trace_stack_spec(nodes, tris, verts, ents, 1, rays, variant=v, flags=flags)
            assert np.array_equal(a["t"], refa["t"]) and np.array_equal(a["entity"], refa["entity"])
    assert (ref["t"] > 0).mean() > 0.2


@pytest.mark.parametrize("kind", ["cycle", "bad_child", "overflow"])
def test_guards(crt, orc, kind):
    v, tris = single_triangle_scene()
    nodes = crafted_stack_chain(70) if kind == "overflow" else crafted_stack(kind)
    ent = orc.identity_entity(nodes.size)
    rays = random_rays(300, [-0.5, -0.5, -1.0], [1.0, 1.0, -0.5], seed=4, toward=[0.3, 0.3, 0.0])
    want = {"cycle": 0x01, "bad_child": 0x04, "overflow": 0x02}[kind]
    for q in (0, 1):
        ref = orc.trace(1, nodes, tris, v, ent, q, rays)
        if q == 0:
            assert ((ref["status"] >> 24) == want).all(), "the crafted buffer does not trip its guard"
        for var in orc.SBVH2_PRESETS:
            h = orc.trace_stack_spec(nodes, tris, v, ent, q, rays, variant=var)
            hf, rf = h["status"] >> 24, ref["status"] >> 24
            if q == 0:
                assert np.array_equal(hf, rf), f"{kind} v={var}"
                assert np.array_equal(h["status"] & 0xFFFFFF, ref["status"] & 0xFFFFFF)
            else:
                # an any-hit walk of the reference stops at its first hit; the speculative walk finishes its round
                # of node steps first and can meet a guard on the way
                hit = ref["t"] > 0
                assert np.array_equal(hf[~hit], rf[~hit]) and ((hf[hit] & rf[hit]) == rf[hit]).all()
            assert np.array_equal(h["t"], ref["t"])


def test_schedule_limits(crt, orc):
    b, ent, rays = _scene(crt, orc, "cloud")
    with pytest.raises(RuntimeError):
        orc.trace_stack_spec(b.nodes, b.triangles, b.vertices, ent, 0, rays[:64], schedule=(8, 1))
    with pytest.raises(RuntimeError):
        orc.trace_stack_spec(b.nodes, b.triangles, b.vertices, ent, 0, rays[:64], schedule=(8, 17))


def test_planted_mistake_fails(crt, orc):
    """The checks notice a twin that drops right leaf children: planted here by turning every right leaf child of
    a copy of the buffers given to the twin alone into an empty leaf."""
    b, ent, rays = _scene(crt, orc, "cloud")
    ref = orc.trace(1, b.nodes, b.triangles, b.vertices, ent, 0, rays)
    bad = b.nodes.copy()
    rp = bad["right"]["min"][:, 3].view(np.int32)
    leaf = rp != -1
    rp[leaf] = rp[leaf] & ~0xF
    h = orc.trace_stack_spec(bad, b.triangles, b.vertices, ent, 0, rays)
    assert not np.array_equal(h["t"], ref["t"])
