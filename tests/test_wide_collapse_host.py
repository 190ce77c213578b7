"""The device BVH4 / BVH8 collapse logic built for the host (no GPU).

tests/cpp/wide_collapse_host.cpp runs the per-node functions the collapse
kernels call (candela_amd/csrc/wide_collapse.h: root_one_slot, make_kid,
expand_one, assign_one) level by level in the device's order.  It is built
three ways -- under the undefined-behaviour sanitizer with every local set to
a pattern, with every local zeroed, and plainly optimized -- and each build's
layout must equal crt_wide_build's, renumbered breadth-first, on SAH trees
(the reference constructor's stack layout) and on linear BVHs (the device
builder's host twin): a value defined only on some paths would make the
builds disagree.  Trees the collapse must refuse come back with their status
bit and an empty root.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT

CLANG = "/opt/rocm/llvm/bin/clang++"
BUILDS = {
    "ubsan_pattern": ["-O1", "-g", "-fsanitize=undefined", "-fno-sanitize-recover=all",
                      "-ftrivial-auto-var-init=pattern"],
    "zero": ["-O2", "-ftrivial-auto-var-init=zero"],
    "plain": ["-O2"],
}


@pytest.fixture(scope="module")
def exes(tmp_path_factory):
    cxx = CLANG if os.path.exists(CLANG) else shutil.which("clang++")
    if cxx is None:
        pytest.skip("clang++ with -ftrivial-auto-var-init is not installed")
    d = tmp_path_factory.mktemp("wide_collapse_host")
    out = {}
    for name, flags in BUILDS.items():
        exe = str(d / f"collapse_{name}")
        subprocess.run([cxx, "-std=c++17", "-ffp-contract=off", *flags, f"-I{ROOT}/include",
                        f"-I{ROOT}/candela_amd/csrc", os.path.join(ROOT, "tests", "cpp", "wide_collapse_host.cpp"),
                        "-o", exe], check=True)
        out[name] = exe
    return out


def run_host(exe, nodes, ntris, width):
    blob = struct.pack("<qqi", nodes.size, ntris, width) + np.ascontiguousarray(nodes).tobytes()
    res = subprocess.run([exe], input=blob, stdout=subprocess.PIPE, stderr=subprocess.PIPE, check=False)
    assert res.returncode == 0, res.stderr.decode()[-2000:]
    total, status = struct.unpack_from("<qI", res.stdout, 0)
    import candela_amd as crt
    groups = np.frombuffer(res.stdout, crt.WIDE_GROUP_DTYPE, offset=12)
    assert groups.size == total * (width // 4)
    return total, status, groups


def _trees(crt):
    import oracle
    from candela_amd import scenes
    from conftest import concat_meshes
    rng = np.random.default_rng(4)
    out = []
    for name, meshes in (("boxes", scenes.boxes_scene()),
                         ("cloud", scenes.icosphere_cloud(count=60, subdivisions=2, seed=9, box=30.0))):
        b = crt.build_bvh(1, meshes)
        out.append((f"sah_{name}", b.nodes, b.triangles, b.vertices))
    for n in (2, 3, 5, 1000, 20000):
        c = rng.uniform(-50, 50, size=(n, 1, 3))
        p = (c + rng.normal(scale=0.8, size=(n, 3, 3))).astype(np.float32)
        verts = scenes.make_vertices(p.reshape(-1, 3))
        nodes, tris = oracle.build_lbvh(verts, np.arange(3 * n, dtype=np.uint32))
        out.append((f"lbvh_{n}", nodes, tris, verts))
    p = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (120, 1))
    verts = scenes.make_vertices(p)
    nodes, tris = oracle.build_lbvh(verts, np.arange(360, dtype=np.uint32))
    out.append(("lbvh_same", nodes, tris, verts))
    return out


@pytest.mark.parametrize("width", [4, 8])
def test_host_builds_equal_crt_wide_build(crt, orc, exes, width):
    import oracle
    from test_gpu_build import _bfs_renumber
    for name, nodes, tris, verts in _trees(crt):
        ent = oracle.identity_entity(nodes.size)
        h = crt.wide_build(1, nodes, tris, verts, ent, width, 2)
        ref, reached = _bfs_renumber(h.groups, width, int(h.entities["node_offset"][0]))
        for build, exe in exes.items():
            total, status, groups = run_host(exe, nodes, tris.size, width)
            assert status == 0, (name, build, status)
            assert total == reached == h.stats["node_count"], (name, build)
            assert groups.tobytes() == ref.tobytes(), f"{name} width {width}: the {build} build differs"


def test_refused_trees_leave_an_empty_root(crt, orc, exes):
    import oracle
    from candela_amd import scenes
    rng = np.random.default_rng(5)
    p = (rng.uniform(-20, 20, size=(500, 1, 3)) + rng.normal(size=(500, 3, 3))).astype(np.float32)
    verts = scenes.make_vertices(p.reshape(-1, 3))
    # leaf starts counted from a triangle offset: every range past the 500 records is refused
    nodes, _ = oracle.build_lbvh(verts, np.arange(1500, dtype=np.uint32), t_offset=100)
    bad_link = oracle.build_lbvh(verts, np.arange(1500, dtype=np.uint32))[0].copy()
    bad_link[0]["left"]["max"][3] = np.array([bad_link.size + 5], np.int32).view(np.float32)[0]
    two = oracle.build_lbvh(scenes.make_vertices(p[:2].reshape(-1, 3)), np.arange(6, dtype=np.uint32), t_offset=7)[0]
    for label, tree, ntris, bit in (("t_offset", nodes, 500, 16), ("bad link", bad_link, 500, 1),
                                    ("two-triangle root at t_offset", two, 2, 16)):
        for width in (4, 8):
            for build, exe in exes.items():
                total, status, groups = run_host(exe, tree, ntris, width)
                assert status & bit, (label, build, status)
                root = groups[:width // 4]
                assert (root["count"] == -1).all(), f"{label}: root not empty ({build})"
