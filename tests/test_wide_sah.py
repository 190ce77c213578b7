"""The SAH collapse of crt_wide_build_ex (CRT_COLLAPSE_SAH), on the host.

* optimality: on small trees every collapse is enumerated by brute force (each
  wide node's slots range over every cover of its subtree by at most `width`
  subtrees, each subtree a leaf where the leaf rule allows or a wide node), and
  the expected cost of the tree the library returns, recomputed from its
  groups, must equal the least enumerated cost;
* structure and traversal: the layout passes the same checks as the greedy
  collapse (tests/test_wide.py) and walks to the brute-force nearest hits;
* errors: an unknown collapse or a non-positive cost is rejected.
"""
import functools

import numpy as np
import pytest

from oracle_cases import random_rays


def _decode_stack(nodes):
    """Binary tree of a stack-layout object: list of dicts (box, leaf range or children), root first."""
    out = [None]

    def child(e, side):
        b = nodes[side][e]
        box = (np.float64(b["min"][:3]), np.float64(b["max"][:3]))
        packed = int(b["min"][3:4].view(np.int32)[0])
        if packed != -1:
            node = {"box": box, "leaf": (packed >> 4, packed & 0xF)}
        else:
            node = {"box": box, "entry": int(b["max"][3:4].view(np.int32)[0])}
        out.append(node)
        return len(out) - 1

    root_lo = np.minimum(nodes["left"]["min"][0, :3], nodes["right"]["min"][0, :3]).astype(np.float64)
    root_hi = np.maximum(nodes["left"]["max"][0, :3], nodes["right"]["max"][0, :3]).astype(np.float64)
    out[0] = {"box": (root_lo, root_hi), "entry": 0}
    work = [0]
    while work:
        i = work.pop()
        e = out[i]["entry"]
        out[i]["kids"] = (child(e, "left"), child(e, "right"))
        for k in out[i]["kids"]:
            if "entry" in out[k]:
                work.append(k)
    # subtree triangle sums and ranges
    for i in range(len(out) - 1, -1, -1):
        n = out[i]
        if "leaf" in n:
            n["tris"], n["first"], n["last"] = n["leaf"][1], n["leaf"][0], n["leaf"][0] + n["leaf"][1]
        else:
            a, b = out[n["kids"][0]], out[n["kids"][1]]
            n["tris"], n["first"], n["last"] = a["tris"] + b["tris"], min(a["first"], b["first"]), max(a["last"],
                                                                                                      b["last"])
    return out


def _half_area(lo, hi):
    d = np.asarray(hi, np.float64) - np.asarray(lo, np.float64)
    return float(d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


def _brute_force_cost(tree, width, max_leaf, nc, tc):
    """Least expected cost over every collapse, by exhaustive enumeration of each wide node's slot covers."""
    def area(i):
        return _half_area(*tree[i]["box"])

    @functools.lru_cache(maxsize=None)
    def covers_open(i):
        """Every cover of node i's subtree with i opened, of at most `width` subtrees."""
        a, b = tree[i]["kids"]
        return [x + y for x in covers_any(a) for y in covers_any(b) if len(x) + len(y) <= width]

    @functools.lru_cache(maxsize=None)
    def covers_any(i):
        return [(i,)] + ([] if "leaf" in tree[i] else covers_open(i))

    @functools.lru_cache(maxsize=None)
    def one_slot(i):
        n = tree[i]
        if "leaf" in n:
            return area(i) * tc * n["tris"]
        options = [area(i) * nc + min(sum(one_slot(k) for k in c) for c in covers_open(i))]
        if n["tris"] <= max_leaf and n["last"] - n["first"] == n["tris"]:
            options.append(area(i) * tc * n["tris"])
        return min(options)

    return area(0) * nc + min(sum(one_slot(k) for k in c) for c in covers_open(0))


def _layout_cost(w, root_box, nc, tc):
    """Expected cost of a built layout: every wide node at its box (the root at the BVH2 root box), every leaf
    slot at its box times its triangles."""
    g = w.groups
    per = w.width // 4
    n = g.size // per
    lo = np.stack([g["lo_x"], g["lo_y"], g["lo_z"]], axis=-1).reshape(n, w.width, 3)
    hi = np.stack([g["hi_x"], g["hi_y"], g["hi_z"]], axis=-1).reshape(n, w.width, 3)
    child, count = g["child"].reshape(n, w.width), g["count"].reshape(n, w.width)
    cost = _half_area(*root_box) * nc
    stack = [0]
    while stack:
        k = stack.pop()
        for j in range(w.width):
            if count[k, j] > 0:
                cost += _half_area(lo[k, j], hi[k, j]) * tc * int(count[k, j])
            elif count[k, j] == 0:
                cost += _half_area(lo[k, j], hi[k, j]) * nc
                stack.append(int(child[k, j]))
    return cost


def _object(crt, meshes):
    import oracle
    b = crt.build_bvh(crt.FORMAT_STACK, meshes)
    return b, oracle.identity_entity(b.nodes.size)


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("max_leaf,nc,tc", [(2, 1.0, 0.5), (4, 1.0, 0.3), (6, 1.0, 1.5), (4, 2.0, 0.1)])
@pytest.mark.parametrize("seed", [1, 2])
def test_sah_collapse_is_least_cost(crt, width, max_leaf, nc, tc, seed):
    from candela_amd import scenes
    meshes = scenes.icosphere_cloud(count=4, subdivisions=1, seed=seed, box=6.0)   # 320 triangles
    b, ent = _object(crt, meshes)
    tree = _decode_stack(b.nodes)
    w = crt.wide_build(1, b.nodes, b.triangles, b.vertices, ent, width, max_leaf, crt.COLLAPSE_SAH, nc, tc)
    best = _brute_force_cost(tree, width, max_leaf, nc, tc)
    got = _layout_cost(w, tree[0]["box"], nc, tc)
    assert abs(got - best) <= 1e-5 * best, (got, best)
    greedy = crt.wide_build(1, b.nodes, b.triangles, b.vertices, ent, width, max_leaf)
    assert _layout_cost(greedy, tree[0]["box"], nc, tc) >= best * (1 - 1e-6)


def test_planted_cost_mistake_fails(crt):
    """The optimality check notices a collapse that is not the cheapest: the greedy collapse, priced by the SAH
    cost with a costly node visit, is dearer than the enumerated least cost."""
    from candela_amd import scenes
    b, ent = _object(crt, scenes.icosphere_cloud(count=4, subdivisions=1, seed=1, box=6.0))
    tree = _decode_stack(b.nodes)
    greedy = crt.wide_build(1, b.nodes, b.triangles, b.vertices, ent, 4, 4)
    best = _brute_force_cost(tree, 4, 4, 1.0, 0.3)
    assert _layout_cost(greedy, tree[0]["box"], 1.0, 0.3) > best * (1 + 1e-4)


@pytest.mark.parametrize("width,max_leaf", [(4, 2), (8, 2), (4, 4)])
def test_sah_layout_walks_to_brute_force_hits(crt, orc, width, max_leaf):
    from candela_amd import scenes
    r = crt.RayIntersector(0)
    r.add_object(0, scenes.boxes_scene())
    hb = r.host_buffers()
    import oracle
    ent = oracle.identity_entity(hb["nodes"].size)
    w = crt.wide_build(0, hb["nodes"], hb["triangles"], hb["vertices"], ent, width, max_leaf, crt.COLLAPSE_SAH)
    g = crt.wide_build(0, hb["nodes"], hb["triangles"], hb["vertices"], ent, width, max_leaf)
    rays = np.concatenate([random_rays(3000, [-4, 0.2, -4], [20, 8, 20], seed=3),
                           random_rays(2000, [-10, 5, -10], [-5, 9, -5], seed=4, toward=[8, 1, 8])])
    h = orc.trace_wide(width, w.groups, w.records, w.entities, 0, rays)
    hg = orc.trace_wide(width, g.groups, g.records, g.entities, 0, rays)
    bt, _ = orc.bruteforce_closest(hb["triangles"], hb["vertices"], ent, rays)
    assert (h["status"] & 0xFF000000 == 0).all()
    assert np.array_equal(h["t"], bt) and np.array_equal(h["t"], hg["t"])
    a = orc.trace_wide(width, w.groups, w.records, w.entities, 1, rays)
    assert np.array_equal(a["t"] > 0, h["t"] > 0)


def test_sah_errors(crt):
    from candela_amd import scenes
    b, ent = _object(crt, scenes.icosphere_cloud(count=2, subdivisions=1, seed=1, box=6.0))
    args = (1, b.nodes, b.triangles, b.vertices, ent, 4, 2)
    with pytest.raises(crt.CrtError, match="collapse"):
        crt.wide_build(*args, 7)
    for nc, tc in ((-1.0, 0.5), (1.0, -0.5), (float("inf"), 0.5), (1.0, float("nan"))):
        with pytest.raises(crt.CrtError, match="cost"):
            crt.wide_build(*args, crt.COLLAPSE_SAH, nc, tc)
    import ctypes as C
    h = C.c_void_p()
    rc = crt.lib().crt_wide_build_ex(1, crt.ptr(b.nodes), b.nodes.size, crt.ptr(b.triangles), b.triangles.size,
                                     crt.ptr(b.vertices), b.vertices.size, crt.ptr(ent), ent.size, None, C.byref(h))
    assert rc == -1 and b"options" in crt.lib().crt_last_error() and not h.value   # CRT_E_INVALID
    with pytest.raises(crt.CrtError, match="width"):
        crt.wide_build(1, b.nodes, b.triangles, b.vertices, ent, 5, 2, crt.COLLAPSE_SAH)
