"""The speculative walks' records do not depend on which rays share a wave, checked on the host twins.

In a round every lane that can step takes a node step, for up to INNER iterations; a lane that cannot step (queue
full, walk over) cannot become able again before the leaf phase.  So each lane takes INNER steps, or fewer when its
own queue fills or its walk ends, then its leaves are tested; the leaf phase's minimum per owner covers only that
owner's tests, and walk ends and refills happen between rounds, delaying a lane without changing its sequence.  A
ray's record -- t, triangle, barycentrics, visits, status bits, triangle tests -- is therefore the same in any wave.
This is what lets the refilling launches (kernel 5), whose waves take rays in run-dependent order, be checked bit for
bit against the grid twin run with the refilling preset's schedule.

Checked here by regrouping: the same rays traced after 13 other rays (every wave boundary moves) and in a random
permutation give byte-identical records, for the wide walk (BVH4, BVH8) and both BVH2 walks, both queries and every
refilling schedule.
"""
import numpy as np
import pytest

from oracle_cases import axis_aligned_rays, random_rays


def _regroup_equal(fn, rays):
    base, bt = fn(rays)
    pad = random_rays(13, [-3, -3, -3], [33, 33, 33], seed=444)
    shifted, st = fn(np.concatenate([pad, rays]))
    perm = np.random.default_rng(5).permutation(rays.size)
    permuted, pt = fn(rays[perm])
    assert shifted[13:].tobytes() == base.tobytes() and np.array_equal(st[13:], bt)
    back = np.empty_like(permuted)
    back[perm] = permuted
    tb = np.empty_like(pt)
    tb[perm] = pt
    assert back.tobytes() == base.tobytes() and np.array_equal(tb, bt)
    return base


def _rays(q):
    return np.concatenate([random_rays(6000, [-3, -3, -3], [33, 33, 33], seed=91, tmax=3.0 if q else 1e6),
                           axis_aligned_rays(800, [0, 0, 0], [30, 30, 30], seed=92)])


@pytest.fixture(scope="module")
def cloud(crt):
    from candela_amd import scenes
    return scenes.icosphere_cloud(count=120, subdivisions=2, seed=13, box=30.0)


@pytest.mark.parametrize("width", [4, 8])
def test_wide_twin_regroups(crt, orc, cloud, width):
    b = crt.build_bvh(0, cloud)
    ent = orc.identity_entity(b.nodes.size)
    w = crt.wide_build(0, b.nodes, b.triangles, b.vertices, ent, width, 2)
    for q in (0, 1):
        for v, sched in orc.SPEC_P_PRESETS.items():
            h = _regroup_equal(lambda r: orc.trace_wide_spec(width, w.groups, w.records, w.entities, q, r,
                                                             schedule=sched, with_tests=True), _rays(q))
            assert (h["t"] > 0).mean() > 0.05


@pytest.mark.parametrize("fmt", [0, 1])
def test_bvh2_twins_regroup(crt, orc, cloud, fmt):
    b = crt.build_bvh(fmt, cloud)
    ent = orc.identity_entity(b.nodes.size)
    fn = orc.trace_stack_spec if fmt == 1 else orc.trace_stackless_spec
    for q in (0, 1):
        for v, sched in orc.SBVH2_P_PRESETS.items():
            _regroup_equal(lambda r: fn(b.nodes, b.triangles, b.vertices, ent, q, r, schedule=sched,
                                        with_tests=True), _rays(q))


def test_regroup_check_notices_a_schedule(crt, orc, cloud):
    """The comparison is sharp enough to see a different schedule: the twin with another (INNER, Q) gives other
    visit counts on some rays."""
    b = crt.build_bvh(0, cloud)
    ent = orc.identity_entity(b.nodes.size)
    rays = _rays(0)
    a = orc.trace_stackless_spec(b.nodes, b.triangles, b.vertices, ent, 0, rays, schedule=(16, 8))
    c = orc.trace_stackless_spec(b.nodes, b.triangles, b.vertices, ent, 0, rays, schedule=(2, 2))
    assert np.array_equal(a["t"], c["t"])
    assert not np.array_equal(a["status"], c["status"])
