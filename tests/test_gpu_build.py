"""GPU parity of the device BVH builder (crt_build_bvh_gpu) against its host twin.

The builder's arithmetic is min / max, one FP32 centroid and one FP32
quantization per triangle, and integer work, so the node and triangle arrays
must equal oracle/oracle_lbvh.c byte for byte: on triangle soups of 2, 3,
1,000 and 200,000 triangles, the boxes scene with mesh ids, an icosphere
cloud, 120 identical triangles (every Morton code equal) and a flat scene.
A second build into the same buffers after the vertices move must match the
twin again (the workspace and its flags are reused).  Rays traced on the GPU
over the built tree must equal the host stack walk over the same buffers.
"""
import sys

import numpy as np
import pytest

from oracle_cases import random_rays
from test_gpu_trace import compare, gpu_trace

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu(crt):
    if crt.device_count() < 1:
        pytest.fail("no HIP device visible")
    return crt


def _soup(n, seed, flat=False):
    from candela_amd import scenes
    rng = np.random.default_rng(seed)
    c = rng.uniform(-50, 50, size=(n, 1, 3))
    p = (c + rng.normal(scale=0.8, size=(n, 3, 3))).astype(np.float32)
    if flat:
        p[:, :, 1] = 0.5
    return scenes.make_vertices(p.reshape(-1, 3)), np.arange(3 * n, dtype=np.uint32), None


def _scene(name):
    from candela_amd import scenes
    from conftest import concat_meshes
    if name == "boxes":
        return concat_meshes(scenes.boxes_scene())
    if name == "cloud":
        return concat_meshes(scenes.icosphere_cloud(count=200, subdivisions=3, seed=5, box=60.0))
    if name == "same":
        p = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (120, 1))
        return scenes.make_vertices(p), np.arange(360, dtype=np.uint32), None
    if name == "flat":
        return _soup(5000, 4, flat=True)
    return _soup(int(name), seed=int(name))


@pytest.mark.parametrize("name", ["2", "3", "1000", "200000", "boxes", "cloud", "same", "flat"])
def test_build_matches_twin(gpu, orc, name):
    log(f"[case] gpu build {name}")
    verts, idx, mids = _scene(name)
    b = gpu.GpuBVH(verts, idx, mids)
    gpu.lib().crt_device_sync()
    on, ot = orc.build_lbvh(verts, idx, mids)
    assert b.nodes.download()[:on.size].tobytes() == on.tobytes()
    assert b.triangles.download().tobytes() == ot.tobytes()
    # move the vertices and build again into the same buffers
    moved = verts.copy()
    moved["position"][:, :3] = moved["position"][:, :3] * np.float32(1.5) + np.float32(3.0)
    b.vertices.upload(moved)
    b.build()
    gpu.lib().crt_device_sync()
    on2, ot2 = orc.build_lbvh(moved, idx, mids)
    assert b.nodes.download()[:on2.size].tobytes() == on2.tobytes()
    assert b.triangles.download().tobytes() == ot2.tobytes()
    log(f"[parity] gpu build {name}: {idx.size // 3} triangles, nodes and triangles byte-identical (two builds)")


def test_trace_built_tree(gpu, orc):
    import oracle
    log("[case] trace over the gpu-built tree")
    verts, idx, mids = _scene("cloud")
    b = gpu.GpuBVH(verts, idx, mids)
    gpu.lib().crt_device_sync()
    nodes, tris = b.nodes.download(), b.triangles.download()
    ent = oracle.identity_entity(nodes.size)
    rays = random_rays(60000, [-2, -2, -2], [62, 62, 62], seed=6)
    for q in (0, 1):
        g = gpu_trace(gpu, b.scene, q, rays, kernel=1)
        o = orc.trace(1, nodes, tris, verts, ent, q, rays)
        compare(g, o, f"gpu-built stack tree q={q}")
        assert g.tobytes() == o.tobytes()


def test_build_errors(gpu):
    import ctypes as C
    verts, idx, _ = _soup(10, 1)
    b = gpu.GpuBVH(verts, idx)
    lib = gpu.lib()
    args = [b.vertices.ptr, b.vertices.count, b.indices.ptr, b.ntris, None, 0, 0, b.nodes.ptr, b.triangles.ptr]
    with pytest.raises(gpu.CrtError):           # workspace too small
        gpu.check(lib.crt_build_bvh_gpu(*args, b.workspace.ptr, 16, None))
    with pytest.raises(gpu.CrtError):           # one triangle
        gpu.check(lib.crt_build_bvh_gpu(b.vertices.ptr, 3, b.indices.ptr, 1, None, 0, 0, b.nodes.ptr,
                                        b.triangles.ptr, b.workspace.ptr, b.workspace.nbytes, None))
    with pytest.raises(gpu.CrtError):           # leaf starts past 2^28
        gpu.check(lib.crt_build_bvh_gpu(b.vertices.ptr, b.vertices.count, b.indices.ptr, b.ntris, None, 0,
                                        (1 << 28) - 5, b.nodes.ptr, b.triangles.ptr, b.workspace.ptr,
                                        b.workspace.nbytes, None))
    assert lib.crt_build_bvh_gpu_workspace(1) == 0


def _bfs_renumber(groups, width, root=0):
    """The wide layout renumbered breadth-first from `root` (children in slot order), as the device collapse
    numbers it."""
    import numpy as np
    per = width // 4
    child = groups["child"].reshape(-1, width)
    count = groups["count"].reshape(-1, width)
    order = [root]
    for node in order:
        for s in range(width):
            if count[node, s] == 0:
                order.append(int(child[node, s]))
    new_of = {old: new for new, old in enumerate(order)}
    out = np.concatenate([groups[old * per:(old + 1) * per] for old in order])
    c = out["child"].reshape(-1, width)
    k = out["count"].reshape(-1, width)
    for i in range(c.shape[0]):
        for s in range(width):
            if k[i, s] == 0:
                c[i, s] = new_of[int(c[i, s])]
    out["child"] = c.reshape(out["child"].shape)
    return out, len(order)


@pytest.mark.parametrize("name", ["2", "3", "1000", "200000", "boxes", "cloud", "same", "flat"])
def test_wide_collapse_on_device(gpu, orc, name):
    """crt_wide_build_gpu over the device-built tree: node count, groups (renumbered breadth-first) and records
    byte-identical to crt_wide_build on the host over the same buffers; walks of both layouts give identical hit
    records and triangle-test counts."""
    import oracle
    log(f"[case] device wide collapse {name}")
    verts, idx, mids = _scene(name)
    b = gpu.GpuBVH(verts, idx, mids)
    gpu.lib().crt_device_sync()
    nodes, tris = b.nodes.download(), b.triangles.download()
    ent = oracle.identity_entity(nodes.size)
    for width in (4, 8):
        w = gpu.GpuWide(b, width)
        h = gpu.wide_build(1, nodes, tris, verts, ent, width, 2)
        assert w.node_count == h.stats["node_count"]
        ref, reached = _bfs_renumber(h.groups, width, int(h.entities["node_offset"][0]))
        assert reached == w.node_count
        got = w.groups.download()[:w.node_count * (width // 4)]
        assert got.tobytes() == ref.tobytes(), f"groups differ ({name}, width {width})"
        assert w.records.download().tobytes() == h.records.tobytes()
        # the same visit order in both layouts: ties resolve alike
        lo, hi = ([-60, -60, -60], [60, 60, 60]) if name.isdigit() or name == "flat" else ([-4, -4, -4], [64, 64, 64])
        rays = random_rays(20000, lo, hi, seed=31)
        for q in (0, 1):
            g = gpu_trace(gpu, w.scene, q, rays, kernel=3, variant=0)
            o = orc.trace_wide(width, h.groups, h.records, h.entities, q, rays)
            compare(g, o, f"device collapse {name} w{width} q={q}")
            assert g.tobytes() == o.tobytes()
        log(f"[parity] device wide collapse {name} width {width}: {w.node_count} nodes byte-identical (breadth-first)")


def test_wide_collapse_errors(gpu):
    import ctypes as C
    verts, idx, _ = _soup(10, 1)
    b = gpu.GpuBVH(verts, idx)
    lib = gpu.lib()
    w = gpu.GpuWide(b, 4)
    n = C.c_int64()
    args = (b.nodes.ptr, b.nodes.count, b.triangles.ptr, b.ntris, b.vertices.ptr, b.vertices.count)
    with pytest.raises(gpu.CrtError):           # width
        gpu.check(lib.crt_wide_build_gpu(*args, 6, w.groups.ptr, w.records.ptr, C.byref(n), w.workspace.ptr,
                                         w.workspace.nbytes, None))
    with pytest.raises(gpu.CrtError):           # workspace too small
        gpu.check(lib.crt_wide_build_gpu(*args, 4, w.groups.ptr, w.records.ptr, C.byref(n), w.workspace.ptr, 64,
                                         None))
    with pytest.raises(gpu.CrtError):           # a vertex index past the vertex array
        gpu.check(lib.crt_wide_build_gpu(b.nodes.ptr, b.nodes.count, b.triangles.ptr, b.ntris, b.vertices.ptr, 2, 4,
                                         w.groups.ptr, w.records.ptr, C.byref(n), w.workspace.ptr,
                                         w.workspace.nbytes, None))
    assert lib.crt_wide_build_gpu_workspace(1, 4) == 0 and lib.crt_wide_build_gpu_workspace(100, 5) == 0


@pytest.mark.parametrize("name", ["2", "3", "1000", "200000", "boxes", "cloud", "same", "flat"])
def test_wide_collapse_async(gpu, name):
    """crt_wide_build_gpu_async (64 levels, nothing read back) writes crt_wide_build_gpu's node count, groups and
    records byte for byte, twice over the same workspace; its scene, which spans ntris nodes, traces to the same
    hit records as the exact one."""
    log(f"[case] async wide collapse {name}")
    verts, idx, mids = _scene(name)
    b = gpu.GpuBVH(verts, idx, mids)
    lo, hi = ([-60, -60, -60], [60, 60, 60]) if name.isdigit() or name == "flat" else ([-4, -4, -4], [64, 64, 64])
    rays = random_rays(20000, lo, hi, seed=37)
    for width in (4, 8):
        w = gpu.GpuWide(b, width)
        a = gpu.GpuWide(b, width, max_levels=64)
        for rep in range(2):
            if rep:
                a.groups.fill_bytes(0x7F)
                a.build()
            n, status = a.result()
            assert status == 0, f"status {status:#x} ({name}, width {width})"
            assert n == w.node_count
            per = width // 4
            assert a.groups.download()[:n * per].tobytes() == w.groups.download()[:n * per].tobytes()
            assert a.records.download().tobytes() == w.records.download().tobytes()
        for q in (0, 1):
            ga = gpu_trace(gpu, a.scene, q, rays, kernel=3, variant=0)
            gw = gpu_trace(gpu, w.scene, q, rays, kernel=3, variant=0)
            assert ga.tobytes() == gw.tobytes(), f"async scene traces differ ({name}, width {width}, q={q})"
        log(f"[parity] async wide collapse {name} width {width}: {n} nodes, identical to the synchronous collapse")


def test_wide_collapse_async_errors(gpu):
    import ctypes as C
    lib = gpu.lib()
    verts, idx, _ = _soup(1000, 3)
    b = gpu.GpuBVH(verts, idx)
    a = gpu.GpuWide(b, 4, max_levels=64)
    n_ok, st = a.result()
    assert st == 0 and n_ok == gpu.GpuWide(b, 4).node_count
    args = (b.nodes.ptr, b.nodes.count, b.triangles.ptr, b.ntris, b.vertices.ptr, b.vertices.count)

    def call(*mid, nodes=None, nverts=None, ws=None):
        a0 = list(args)
        if nodes is not None:
            a0[0] = nodes
        if nverts is not None:
            a0[5] = nverts
        return lib.crt_wide_build_gpu_async(*a0, *mid, a.groups.ptr, a.records.ptr, a.out.ptr, a.out.ptr + 16,
                                            a.workspace.ptr, a.workspace.nbytes if ws is None else ws, None)

    for bad in ((6, 64), (4, 0), (4, gpu.WIDE_GPU_MAX_LEVELS + 1)):
        with pytest.raises(gpu.CrtError):
            gpu.check(call(*bad))
    with pytest.raises(gpu.CrtError):           # workspace too small
        gpu.check(call(4, 64, ws=64))
    assert lib.crt_wide_build_gpu_async_workspace(1, 4, 8) == 0
    assert lib.crt_wide_build_gpu_async_workspace(100, 4, 0) == 0
    assert lib.crt_wide_build_gpu_async(None, 1, None, 2, None, 1, 4, 8, None, None, None, None, None, 0, None) < 0
    # data errors come back in the status word
    gpu.check(call(4, 64, nverts=2))
    assert a.result()[1] & gpu.WIDE_GPU_BAD_VERTEX
    gpu.check(call(4, 64))
    assert a.result()[1] == 0
    need = a.levels
    assert 2 <= need < 64
    gpu.check(call(4, need))                    # exactly the levels the tree needs
    assert a.result() == (n_ok, 0)
    gpu.check(call(4, need - 1))
    assert a.result()[1] == gpu.WIDE_GPU_TOO_DEEP
    gpu.check(call(4, 1))
    assert a.result()[1] == gpu.WIDE_GPU_TOO_DEEP
    nodes = b.nodes.download()
    assert nodes[0]["left"]["min"][3].view(np.int32) == -1          # the root's left child is inner
    nodes[0]["left"]["max"][3] = np.array([b.nodes.count + 5], np.int32).view(np.float32)[0]
    bad_nodes = gpu.DeviceArray.from_host(nodes)
    gpu.check(call(4, 64, nodes=bad_nodes.ptr))
    assert a.result()[1] & gpu.WIDE_GPU_BAD_LINK
    gpu.check(call(4, 64))
    assert a.result()[1] == 0


def test_wide_collapse_error_layouts_trace_safely(gpu):
    """Every layout the collapse flags is safe to trace: a tree built with a triangle offset (leaf starts past the
    records), a truncated collapse (max_levels one short) and a bad link each set their status bit, and closest-
    and any-hit traces launched right after, over the scene that spans ntris nodes, end normally with no hit and no
    guard bit (the root is written empty)."""
    verts, idx, _ = _soup(3000, 12)
    rays = random_rays(5000, [-60, -60, -60], [60, 60, 60], seed=43)

    def traced_empty(a, what):
        for q in (0, 1):
            for kernel, variant in ((3, 0), (2, 8), (1, 0)):
                h = gpu_trace(gpu, a.scene, q, rays, kernel=kernel, variant=variant)
                assert (h["t"] == -1).all() and (h["tri"] == -1).all(), f"{what}: hits after an error (q={q})"
                assert (h["status"] & 0xFF000000 == 0).all(), f"{what}: guard bits after an error"

    log("[case] collapse of a tree built with t_offset 100")
    off = gpu.GpuBVH(verts, idx, t_offset=100)
    with pytest.raises(gpu.CrtError, match="leaf range"):
        gpu.GpuWide(off, 4)
    for width in (4, 8):
        a = gpu.GpuWide(off, width, max_levels=64)
        n, status = a.result()
        assert status & gpu.WIDE_GPU_BAD_LEAF, f"status {status:#x}"
        traced_empty(a, f"t_offset width {width}")

    log("[case] truncated collapse")
    b = gpu.GpuBVH(verts, idx)
    for width in (4, 8):
        full = gpu.GpuWide(b, width, max_levels=64)
        assert full.result()[1] == 0
        need = full.levels                 # the levels this width's tree needs
        assert need >= 2
        a = gpu.GpuWide(b, width, max_levels=need - 1)
        a.groups.fill_bytes(0x3C)        # stale memory in the levels that are never expanded
        a.build()
        assert a.result()[1] & gpu.WIDE_GPU_TOO_DEEP
        traced_empty(a, f"truncated width {width}")

    log("[case] bad link")
    nodes = b.nodes.download()
    nodes[0]["left"]["max"][3] = np.array([b.nodes.count + 5], np.int32).view(np.float32)[0]
    b.nodes.upload(nodes)
    a = gpu.GpuWide(b, 4, max_levels=64)
    assert a.result()[1] & gpu.WIDE_GPU_BAD_LINK
    traced_empty(a, "bad link")
    log("[parity] error layouts trace to misses")


def test_rebuild_graph(gpu):
    """The per-frame rebuild in one graph: crt_build_bvh_gpu, crt_wide_build_gpu_async and a closest-hit and an
    any-hit trace over the new layout, captured once and replayed on moving geometry (eager uploads and
    synchronizations between replays).  Every replay gives the eager synchronous pipeline's hit records."""
    from candela_amd import scenes
    from conftest import concat_meshes
    verts, idx, mids = concat_meshes(scenes.icosphere_cloud(count=150, subdivisions=3, seed=8, box=50.0))
    rays = random_rays(30000, [-4, -4, -4], [54, 54, 54], seed=41)
    s = gpu.Stream()
    b = gpu.GpuBVH(verts, idx, mids, stream=s)
    a = gpu.GpuWide(b, 4, stream=s, max_levels=64)
    d_rays = gpu.DeviceArray.from_host(rays)
    hits = [gpu.DeviceArray(gpu.HIT_DTYPE, rays.size) for _ in range(2)]
    for q in (0, 1):   # eager warm-up
        gpu.trace(a.scene, q, d_rays, hits[q], kernel=3, variant=0, stream=s)
    s.sync()
    g = gpu.Graph(s)
    with g:
        b.build(s)
        a.build(s)
        for q in (0, 1):
            gpu.trace(a.scene, q, d_rays, hits[q], kernel=3, variant=0, stream=s)
    assert g.handle
    rng = np.random.default_rng(9)
    ref_b = gpu.GpuBVH(verts, idx, mids)
    for frame in range(4):
        moved = verts.copy()
        moved["position"][:, :3] += rng.normal(scale=0.3, size=(verts.size, 3)).astype(np.float32)
        b.vertices.upload(moved)
        for h in hits:
            h.fill_bytes(0xFF)
        s.sync()
        g.launch()
        s.sync()
        got = [h.download() for h in hits]
        n, status = a.result()
        # the eager synchronous pipeline on the same vertices
        ref_b.vertices.upload(moved)
        ref_b.build()
        w = gpu.GpuWide(ref_b, 4)
        assert (n, status) == (w.node_count, 0)
        for q in (0, 1):
            ref = gpu_trace(gpu, w.scene, q, rays, kernel=3, variant=0)
            assert got[q].tobytes() == ref.tobytes(), f"replay {frame} query {q} differs from the eager pipeline"
        log(f"[parity] rebuild graph replay {frame}: {n} wide nodes, hit records identical to the eager pipeline")
    g.close()
    s.close()
