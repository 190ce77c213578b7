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
