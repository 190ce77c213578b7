"""Shared scene / ray builders for the oracle and GPU parity tests."""
import numpy as np

RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmax", "<f4"), ("dir", "<f4", 3), ("id", "<u4")])


def random_rays(n: int, lo, hi, seed: int, tmax: float = 1e6, toward=None) -> np.ndarray:
    """Rays with origins uniform in the box [lo, hi] and uniform directions (or aimed at `toward` +- noise)."""
    rng = np.random.default_rng(seed)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    o = rng.uniform(lo, hi, size=(n, 3))
    if toward is None:
        d = rng.normal(size=(n, 3))
    else:
        d = np.asarray(toward, np.float64)[None, :] + rng.normal(scale=0.35 * np.linalg.norm(hi - lo), size=(n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros(n, RAY_DTYPE)
    r["org"] = o.astype(np.float32)
    r["dir"] = d.astype(np.float32)
    r["tmax"] = np.float32(tmax)
    r["id"] = np.arange(n, dtype=np.uint32)
    return r


def axis_aligned_rays(n: int, lo, hi, seed: int) -> np.ndarray:
    """Rays along +-x/y/z (inverse direction has infinities)."""
    rng = np.random.default_rng(seed)
    r = np.zeros(n, RAY_DTYPE)
    r["org"] = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    axis = rng.integers(0, 3, n)
    sign = rng.choice([-1.0, 1.0], n)
    d = np.zeros((n, 3), np.float32)
    d[np.arange(n), axis] = sign
    r["dir"] = d
    r["tmax"] = np.float32(1e6)
    r["id"] = np.arange(n, dtype=np.uint32)
    return r


def _fbits(i: int) -> np.float32:
    return np.array([i], np.int32).view(np.float32)[0]


def single_triangle_scene():
    """One triangle a=(0,0,0), b=(1,0,0), c=(0,1,0) in the z=0 plane."""
    from candela_amd.scenes import make_vertices
    v = make_vertices(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32))
    tris = np.zeros(1, np.dtype([("packed", "<i4", 4)]))
    tris["packed"][0] = [0, 1, 2, 7]
    return v, tris


def crafted_stackless(kind: str):
    """Small hand-written stackless buffers that exercise one traversal guard each.

    Node 0 is an inner root; leaves reference the single triangle 0 (valid).
      'bad_link'  leaf 2's miss link points far outside the entity -> BAD_INDEX
      'cycle'     leaf 1's miss link points back to itself -> ITER_CAP at 1024
    """
    nd = np.dtype([("min", "<f4", 4), ("max", "<f4", 4)])
    n = np.zeros(3, nd)
    for k in range(3):
        n["min"][k, :3] = [-1, -1, -1]
        n["max"][k, :3] = [2, 2, 2]
    n["min"][0, 3] = _fbits(-1)
    n["max"][0, 3] = _fbits(-1)
    leaf = (0 << 4) | 1
    n["min"][1, 3] = _fbits(leaf)
    n["min"][2, 3] = _fbits(leaf)
    if kind == "bad_link":
        n["max"][1, 3] = _fbits(2)
        n["max"][2, 3] = _fbits(1000)
    elif kind == "cycle":
        n["max"][1, 3] = _fbits(1)
        n["max"][2, 3] = _fbits(-1)
    else:
        raise ValueError(kind)
    return n


def crafted_stack_chain(depth: int = 70):
    """Stack buffer whose every entry has two inner children that the ray enters:
    entry k -> (k+1, last).  Each step pushes one entry, so the walk aborts with
    STACK_OVERFLOW once 63 entries are pushed (TraverseBVHStack.glsl:293-295).
    The last entry holds two leaves of triangle 0."""
    bd = np.dtype([("min", "<f4", 4), ("max", "<f4", 4)])
    sd = np.dtype([("left", bd), ("right", bd)])
    n = np.zeros(depth, sd)
    for side in ("left", "right"):
        n[side]["min"][:, :3] = [-1, -1, -1]
        n[side]["max"][:, :3] = [2, 2, 2]
    last = depth - 1
    for k in range(depth - 1):
        n["left"]["min"][k, 3] = _fbits(-1)
        n["right"]["min"][k, 3] = _fbits(-1)
        n["left"]["max"][k, 3] = _fbits(k + 1)
        n["right"]["max"][k, 3] = _fbits(last)
    n["left"]["min"][last, 3] = _fbits(1)
    n["right"]["min"][last, 3] = _fbits(1)
    return n


WIDE_KINDS = ("cycle", "cycle_leaf", "fan", "bad_child", "bad_record")
WIDE_MAX_RECORD = 1 << 26   # leaf starts the speculative walk's queue entries can pack (start << 4 | count)


def crafted_wide(kind: str, width: int):
    """Hand-written BVH4 / BVH8 buffers (crt_wide_group layout) that exercise one traversal guard each.

    Every used slot has the box [0,1]^3; three one-triangle records lie in the planes x = 0.5 (record 0,
    triangle 0), x = 0.25 (record 1, triangle 1) and x = 0.75 (record 2, triangle 2), covering the box's
    y/z extent.  Returns (groups, records, entity node count).
      'cycle'       root slot 0 is an inner child pointing at the root -> ITER_CAP after 1024 visits
      'cycle_leaf'  as 'cycle' plus a leaf of record 0 in slot 1 -> the leaf is tested on the way
                    (the sequential walk pushes the inner child, tests the leaf and pops it each visit;
                    the speculative walk queues it each visit), then ITER_CAP
      'fan'         every slot an inner child pointing at the root -> STACK_OVERFLOW after
                    capacity // (width - 1) + 1 visits
      'bad_child'   leaf of record 0, an inner child outside the entity's node range (node 1 exists in
                    the buffer), one at -5 (an empty slot: inner children need an index >= 0), one far
                    outside the buffer; BVH8: a leaf of record 1 in the second group -> BAD_INDEX,
                    leaves still tested
      'bad_record'  leaves whose start is negative or too large to pack beside their count, next to a
                    leaf of record 1 -> BAD_INDEX from the speculative walk, which never reads them;
                    the sequential walks read leaf records unchecked, so they must not be given it
    """
    from oracle import TRI_RECORD_DTYPE, WIDE_GROUP_DTYPE
    if kind not in WIDE_KINDS:
        raise ValueError(kind)
    per = width // 4
    nodes = 2 if kind == "bad_child" else 1
    g = np.zeros(nodes * per, WIDE_GROUP_DTYPE)
    g["child"] = -1
    g["count"] = -1
    for f in ("lo_x", "lo_y", "lo_z"):
        g[f] = 0.0
    for f in ("hi_x", "hi_y", "hi_z"):
        g[f] = 1.0
    slots = [(0, k) for k in range(4)] + ([(1, k) for k in range(4)] if width == 8 else [])

    def put(i, child, count):
        grp, k = slots[i]
        g["child"][grp, k] = child
        g["count"][grp, k] = count

    if kind == "cycle":
        put(0, 0, 0)
    elif kind == "cycle_leaf":
        put(0, 0, 0)
        put(1, 0, 1)
    elif kind == "fan":
        for i in range(width):
            put(i, 0, 0)
    elif kind == "bad_child":
        put(0, 0, 1)
        put(1, 1, 0)
        put(2, -5, 0)
        put(3, 1000, 0)
        if width == 8:
            put(4, 1, 1)
        g["child"][per] = [2, -1, -1, -1]     # node 1: a leaf of record 2 (outside the entity; never walked)
        g["count"][per] = [1, -1, -1, -1]
    else:
        put(0, -3, 1)
        put(1, WIDE_MAX_RECORD, 1)
        put(2, 1, 1)
        put(3, WIDE_MAX_RECORD - 1, 2)
    r = np.zeros(3, TRI_RECORD_DTYPE)
    for i, x in enumerate((0.5, 0.25, 0.75)):
        r["v0"][i] = [x, -1, -1]
        r["v1"][i] = [x, 3, -1]
        r["v2"][i] = [x, -1, 3]
        r["tri"][i] = i
        r["mesh"][i] = 0
    return g, r, 1


def crafted_wide_rays(seed: int = 5) -> np.ndarray:
    """300 rays for crafted_wide: 200 from x = -1 toward +x (most enter the box), 60 from inside the box in
    random directions, 40 that miss it (y = 5, toward +x)."""
    rng = np.random.default_rng(seed)
    r = np.zeros(300, RAY_DTYPE)
    o = np.empty((300, 3))
    d = np.empty((300, 3))
    o[:200] = np.c_[np.full(200, -1.0), rng.uniform(0.02, 0.98, (200, 2))]
    d[:200] = np.c_[np.ones(200), rng.uniform(-0.2, 0.2, (200, 2))]
    o[200:260] = rng.uniform(0.05, 0.95, (60, 3))
    d[200:260] = rng.normal(size=(60, 3))
    o[260:] = np.c_[np.full(40, -1.0), np.full(40, 5.0), rng.uniform(0, 1, 40)]
    d[260:] = np.c_[np.ones(40), np.zeros(40), np.zeros(40)]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r["org"] = o.astype(np.float32)
    r["dir"] = d.astype(np.float32)
    r["tmax"] = np.float32(1e6)
    r["id"] = np.arange(300, dtype=np.uint32)
    return r


def instanced_bvh2(crt, fmt: int):
    """Two objects (boxes, an icosphere cloud) built in BVH2 format `fmt` (0 stackless, 1 stack) and concatenated
    as AddObject does, with six entities (rotations about y and translations, one of them translucent).  Returns
    (nodes, triangles, vertices, entities)."""
    from candela_amd import scenes
    b0 = crt.build_bvh(fmt, scenes.boxes_scene())
    b3 = crt.build_bvh(fmt, scenes.icosphere_cloud(count=20, subdivisions=2, seed=3, box=16.0),
                       t_offset=b0.triangles.size)
    nodes = np.concatenate([b0.nodes, b3.nodes])
    t3 = b3.triangles.copy()
    t3["packed"][:, :3] += b0.vertices.size
    tris = np.concatenate([b0.triangles, t3])
    verts = np.concatenate([b0.vertices, b3.vertices])
    ents = np.zeros(6, crt.ENTITY_DTYPE)
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
    return nodes, tris, verts, ents


def crafted_stack(kind: str):
    """Small hand-written stack-layout buffers that exercise one traversal guard each; every box is [-1, 2]^3 and
    every leaf holds triangle 0.
      'cycle'      entry 0: left an inner child that is entry 0 again, right a leaf -> ITER_CAP at 1024
      'bad_child'  entry 0: left a leaf, right an inner child at entry 1000 -> BAD_INDEX
    (STACK_OVERFLOW: crafted_stack_chain.)"""
    bd = np.dtype([("min", "<f4", 4), ("max", "<f4", 4)])
    sd = np.dtype([("left", bd), ("right", bd)])
    n = np.zeros(1, sd)
    for side in ("left", "right"):
        n[side]["min"][:, :3] = [-1, -1, -1]
        n[side]["max"][:, :3] = [2, 2, 2]
    leaf = (0 << 4) | 1
    if kind == "cycle":
        n["left"]["min"][0, 3] = _fbits(-1)
        n["left"]["max"][0, 3] = _fbits(0)
        n["right"]["min"][0, 3] = _fbits(leaf)
    elif kind == "bad_child":
        n["left"]["min"][0, 3] = _fbits(leaf)
        n["right"]["min"][0, 3] = _fbits(-1)
        n["right"]["max"][0, 3] = _fbits(1000)
    else:
        raise ValueError(kind)
    return n
