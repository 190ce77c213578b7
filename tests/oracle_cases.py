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
