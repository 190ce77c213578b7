"""Traversal guards of the wide walks' host twins on hand-written BVH4 / BVH8 buffers.

Each crafted buffer (oracle_cases.crafted_wide) trips one guard, and what the
format fixes is checked per ray from the geometry alone, in float64: which
rays enter the root's child box [0,1]^3 (slab test), the iteration cap of
1024 visits, the stack capacity (a visit that would push past it sets
STACK_OVERFLOW, so a fan of width children that all point back at the root
overflows on visit capacity // (width - 1) + 1), the node-range and record
guards (BAD_INDEX on the first visit, the visit's leaves still tested), and
the nearest triangle plane the ray crosses.  Both the sequential twin
(oracle_trace_wide) and the speculative twin (oracle_trace_wide_spec, every
preset) are held to it; the GPU kernels are then held to the twins.
"""
import numpy as np
import pytest

from oracle_cases import WIDE_KINDS, crafted_wide, crafted_wide_rays

ITER_CAP, STACK_OVERFLOW, BAD_INDEX = 1 << 24, 1 << 25, 1 << 26
PLANES = {0: 0.5, 1: 0.25, 2: 0.75}   # record -> plane x of its triangle (the triangle covers the box face)


def _geometry(rays):
    """Per ray: enters the unit box (float64 slab test), clear of the box edges, and t at each plane."""
    o = rays["org"].astype(np.float64)
    d = rays["dir"].astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
        a, b = (0.0 - o) * inv, (1.0 - o) * inv
    near = np.nanmax(np.minimum(a, b), axis=1)
    far = np.nanmin(np.maximum(a, b), axis=1)
    enters = far >= np.maximum(near, 0.0)
    clear = np.abs(far - np.maximum(near, 0.0)) > 1e-4
    planes = {}
    for rec, x in PLANES.items():
        with np.errstate(divide="ignore", invalid="ignore"):
            t = (x - o[:, 0]) / d[:, 0]
        p = o + t[:, None] * d
        inside = (t > 0) & (p[:, 1] >= -1) & (p[:, 2] >= -1) & (p[:, 1] + p[:, 2] <= 2)
        planes[rec] = np.where(inside, t, np.inf)
    return enters, clear, planes


def _expect(kind, width, enters, planes, cap, spec):
    """Expected (flags, visits or None, leaf records that are tested) for rays that enter the box."""
    if kind == "cycle":
        return ITER_CAP, 1024, []
    if kind == "cycle_leaf":
        return ITER_CAP, None, [0]
    if kind == "fan":
        return STACK_OVERFLOW, cap // (width - 1) + 1, []
    if kind == "bad_child":
        return BAD_INDEX, 1, [0, 1] if width == 8 else [0]
    assert spec
    return BAD_INDEX, 1, [1]


def _check(kind, width, q, rays, hits, cap, spec):
    enters, clear, planes = _geometry(rays)
    flags, visits, recs = _expect(kind, width, enters, planes, cap, spec)
    got_flags = hits["status"] & 0xFF000000
    got_vis = hits["status"] & 0x00FFFFFF
    assert clear.mean() > 0.9
    sel = clear & enters
    assert sel.sum() > 100
    assert (got_flags[clear & ~enters] == 0).all() and (got_vis[clear & ~enters] == 1).all()
    assert (hits["t"][clear & ~enters] == -1).all()
    t_exp = np.full(rays.size, np.inf)
    tri_exp = np.full(rays.size, -1)
    for rec in recs:
        nearer = planes[rec] < t_exp
        t_exp = np.where(nearer, planes[rec], t_exp)
        tri_exp = np.where(nearer, rec, tri_exp)
    hit = np.isfinite(t_exp)
    # an any-hit walk ends at its first hit: the sequential walk tests a visit's leaves before the next
    # visit, the speculative walk at the end of its round of node steps
    ended_early = sel & hit if q == 1 else np.zeros(rays.size, bool)
    trips = sel & ~ended_early
    if kind == "bad_record" or (kind == "bad_child" and spec):
        trips = sel                     # the guard trips within the first round, before any leaf test
    assert (got_flags[trips] == flags).all(), f"{kind} w{width} q={q}: guard flags"
    if visits is not None:
        assert (got_vis[trips] == visits).all(), f"{kind} w{width} q={q}: visits {np.unique(got_vis[trips])}"
    assert (got_flags[ended_early & ~trips] == 0).all()
    assert np.array_equal(hits["t"][sel] > 0, hit[sel]), f"{kind} w{width} q={q}: hit / miss"
    h = sel & hit
    if q == 0:
        rel = np.abs(hits["t"][h] - t_exp[h]) / t_exp[h]
        assert (rel < 1e-5).all(), f"{kind} w{width}: nearest t"
        assert np.array_equal(hits["tri"][h], tri_exp[h]), f"{kind} w{width}: nearest triangle"
    else:
        # any hit: the t of one of the tested planes the ray crosses
        near_one = np.zeros(rays.size, bool)
        for rec in recs:
            near_one |= np.abs(hits["t"] - planes[rec]) <= 1e-5 * planes[rec]
        assert near_one[h].all(), f"{kind} w{width}: any-hit t is no tested triangle's"


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("kind", [k for k in WIDE_KINDS if k != "bad_record"])   # read unchecked
def test_sequential_twin_guards(orc, kind, width):
    g, r, nc = crafted_wide(kind, width)
    ent = orc.identity_entity(nc)
    rays = crafted_wide_rays()
    for q in (0, 1):
        for cap in (72, 80):
            hits = orc.trace_wide(width, g, r, ent, q, rays, stack_cap=cap)
            _check(kind, width, q, rays, hits, cap, spec=False)


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("kind", WIDE_KINDS)
def test_speculative_twin_guards(orc, kind, width):
    g, r, nc = crafted_wide(kind, width)
    ent = orc.identity_entity(nc)
    rays = crafted_wide_rays()
    for q in (0, 1):
        for v, (depth, _, _) in orc.SPEC_PRESETS.items():
            hits = orc.trace_wide_spec(width, g, r, ent, q, rays, variant=v)
            _check(kind, width, q, rays, hits, depth + 64, spec=True)


def test_planted_guard_mistake_fails(orc):
    """The check notices a record guard that never trips: planted on the host by turning the two bad leaves
    into valid leaves of record 2."""
    g, r, nc = crafted_wide("bad_record", 4)
    g["child"][0, :2] = 2
    ent = orc.identity_entity(nc)
    rays = crafted_wide_rays()
    with pytest.raises(AssertionError):
        _check("bad_record", 4, 0, rays, orc.trace_wide_spec(4, g, r, ent, 0, rays), 72, spec=True)
