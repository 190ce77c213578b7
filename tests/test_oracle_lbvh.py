"""Pins of the device builder's host twin (oracle/oracle_lbvh.c), no GPU needed.

* the Morton bit spread against a plain bit loop for all 1,024 inputs;
* leaf order is sorted by the Morton codes of the triangle centroids, recomputed
  here in numpy float32 (same quantization), ties kept in input order;
* structure: walking from entry 0 reaches every triangle exactly once; every
  node box is exactly the union of the boxes of the triangles below it; the
  children of every inner node split its leaf range; leaves hold 1 or 2
  triangles; t_offset shifts every leaf start;
* traversal: the BVH4 collapse of the built tree, walked with the wide walk,
  reports the brute-force nearest distance on every ray (the tree differs from
  BuildBVH's, the nearest hit does not); the reference's stack walk, whose box
  test has no margin, agrees to 4 ulps and exactly on 99.9% of the rays;
* degenerate inputs: 2 and 3 triangles, all triangles identical (every code
  equal), a flat scene (one axis of zero extent).
"""
import numpy as np
import pytest

from oracle_cases import random_rays


def spread_ref(v):
    out = 0
    for b in range(10):
        out |= ((v >> b) & 1) << (3 * b)
    return out


def test_spread3(orc):
    for v in range(1024):
        assert orc.spread3(v) == spread_ref(v)


def _soup(n, seed, flat=False):
    from candela_amd import scenes
    rng = np.random.default_rng(seed)
    c = rng.uniform(-20, 20, size=(n, 1, 3))
    p = (c + rng.normal(scale=0.8, size=(n, 3, 3))).astype(np.float32)
    if flat:
        p[:, :, 1] = 0.5
    return scenes.make_vertices(p.reshape(-1, 3)), np.arange(3 * n, dtype=np.uint32)


def _morton_np(verts, idx):
    p = verts["position"][:, :3][idx.reshape(-1, 3)]
    lo, hi = p.min(axis=1), p.max(axis=1)
    c = (lo + hi) * np.float32(0.5)
    clo, chi = c.min(axis=0), c.max(axis=0)
    ext = chi - clo
    with np.errstate(divide="ignore"):
        scale = np.where(ext > 0, np.float32(1) / ext, np.float32(0)).astype(np.float32)
    q = np.clip(((c - clo) * scale) * np.float32(1024), 0, 1023).astype(np.uint32)
    spread = np.vectorize(spread_ref, otypes=[np.uint32])
    return (spread(q[:, 0]) << 2) | (spread(q[:, 1]) << 1) | spread(q[:, 2])


def _walk(nodes, n, t_offset=0):
    """(leaf starts, leaf counts, per-entry reached flag) from entry 0; checks boxes on the way."""
    seen = np.zeros(nodes.size, bool)
    leaves = []
    stack = [0]
    while stack:
        k = stack.pop()
        assert not seen[k]
        seen[k] = True
        for side in ("left", "right"):
            ch = nodes[side][k]
            w = ch["min"][3:4].view(np.int32)[0]
            if w == -1:
                stack.append(int(ch["max"][3:4].view(np.int32)[0]))
            else:
                leaves.append(((w >> 4) - t_offset, w & 0xF))
    return leaves, seen


def _check_structure(crt, orc, verts, idx, t_offset=0):
    n = idx.size // 3
    mesh = np.arange(n, dtype=np.int32) % 7
    nodes, tris = orc.build_lbvh(verts, idx, mesh, 0, t_offset)
    assert nodes.size == n - 1 and tris.size == n
    leaves, seen = _walk(nodes, n, t_offset)
    starts = sorted(s for s, _ in leaves)
    cover = np.concatenate([np.arange(s, s + c) for s, c in sorted(leaves)])
    assert np.array_equal(cover, np.arange(n)), "every triangle exactly once, in leaf order"
    assert all(c in (1, 2) for _, c in leaves)
    assert starts[0] == 0
    # leaf order = sorted Morton order (stable)
    keys = _morton_np(verts, idx)
    tri_of = {tuple(idx.reshape(-1, 3)[t]): t for t in range(n)}
    order = np.array([tri_of[tuple(tris["packed"][k, :3].astype(np.uint32))] for k in range(n)]) \
        if len(tri_of) == n else None
    if order is not None:
        assert np.all(np.diff(keys[order].astype(np.int64)) >= 0)
        same = np.diff(keys[order].astype(np.int64)) == 0
        assert np.all(np.diff(order)[same] > 0), "ties keep input order"
        assert np.array_equal(tris["packed"][:, 3], mesh[order])
    # every box is the exact union of its triangles' boxes
    pos = verts["position"][:, :3]
    tb_lo = pos[tris["packed"][:, :3]].min(axis=1)
    tb_hi = pos[tris["packed"][:, :3]].max(axis=1)

    def union(k):
        lo = np.full(3, np.inf, np.float32)
        hi = np.full(3, -np.inf, np.float32)
        for side in ("left", "right"):
            ch = nodes[side][k]
            w = ch["min"][3:4].view(np.int32)[0]
            if w == -1:
                clo, chi = union(int(ch["max"][3:4].view(np.int32)[0]))
            else:
                s, c = (w >> 4) - t_offset, w & 0xF
                clo, chi = tb_lo[s:s + c].min(0), tb_hi[s:s + c].max(0)
            assert np.array_equal(ch["min"][:3], clo) and np.array_equal(ch["max"][:3], chi)
            lo, hi = np.minimum(lo, clo), np.maximum(hi, chi)
        return lo, hi

    if n <= 4000:
        import sys
        sys.setrecursionlimit(10000)
        union(0)
    return nodes, tris


@pytest.mark.parametrize("n", [2, 3, 17, 500, 3000])
def test_structure_soup(crt, orc, n):
    verts, idx = _soup(n, seed=n)
    _check_structure(crt, orc, verts, idx)
    _check_structure(crt, orc, verts, idx, t_offset=1234)


def test_degenerate(crt, orc):
    from candela_amd import scenes
    p = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (40, 1))
    _check_structure(crt, orc, scenes.make_vertices(p), np.arange(120, dtype=np.uint32))   # all codes equal
    verts, idx = _soup(800, seed=3, flat=True)
    _check_structure(crt, orc, verts, idx)
    with pytest.raises(ValueError):
        orc.build_lbvh(*_soup(1, seed=1))


@pytest.mark.parametrize("scene", ["boxes", "cloud"])
def test_nearest_hit_equals_brute_force(crt, orc, scene):
    import oracle
    from candela_amd import scenes
    from conftest import concat_meshes
    meshes = scenes.boxes_scene() if scene == "boxes" else scenes.icosphere_cloud(count=60, subdivisions=2, seed=8,
                                                                                    box=30.0)
    verts, idx, mids = concat_meshes(meshes)
    nodes, tris = orc.build_lbvh(verts, idx, mids)
    ent = oracle.identity_entity(nodes.size)
    lo, hi = ([-4, 0.2, -4], [20, 8, 20]) if scene == "boxes" else ([-2, -2, -2], [32, 32, 32])
    rays = random_rays(6000, lo, hi, seed=11)
    got = orc.trace(1, nodes, tris, verts, ent, 0, rays)
    ref_t, _ = orc.bruteforce_closest(tris, verts, ent, rays)
    assert (got["status"] & 0xFF000000 == 0).all()
    # The reference's stack walk has no margin in its box test: where a triangle lies in a face of its box
    # (the boxes scene), the box entry can round a few ulps past the triangle's t and the walk then keeps a
    # hit that is an ulp farther.  Its hits agree with brute force to 4 ulps, and exactly on almost every ray.
    hit = ref_t > 0
    assert np.array_equal(got["t"] > 0, hit)
    assert np.all(np.abs(got["t"][hit] - ref_t[hit]) <= 4 * np.spacing(ref_t[hit]))
    assert np.mean(got["t"] == ref_t) > 0.999
    anyh = orc.trace(1, nodes, tris, verts, ent, 1, rays)
    assert np.array_equal(anyh["t"] > 0, hit)
    # the wide walk (its box test keeps a 2^-19 margin) over the collapsed tree is exact
    w = crt.wide_build(1, nodes, tris, verts, ent, 4, 2)
    assert w.stats["record_count"] == tris.size
    wide = orc.trace_wide(4, w.groups, w.records, w.entities, 0, rays)
    assert np.array_equal(wide["t"], ref_t)
