"""Writes the constructor fixtures of tests/golden/ from the reference constructor.

    make -C oracle && python tests/golden/make_golden.py

Runs oracle/_ref/ref_bvh (Source/Core/BVH/BVHConstructor.cpp compiled from the
reference tree, see oracle/Makefile) on seeded scenes and stores
  boxes_bvh.npz     the full stackless / stack node and triangle buffers of the
                    16x16 boxes scene (3,074 triangles)
  bvh_hashes.json   SHA-256 of the buffers of larger scenes
The fixtures are the reference's outputs; libcandela_rt's constructor is not
used here.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle  # noqa: E402
from candela_amd import scenes  # noqa: E402  (scene generators only)
from conftest import concat_meshes  # noqa: E402


def sha(a):
    return hashlib.sha256(a.tobytes()).hexdigest()


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/ref_bvh missing: run make -C oracle with the reference tree present")
    out = {}
    v, i, m = concat_meshes(scenes.boxes_scene())
    sl_n, sl_t, _ = oracle.ref_build(0, v, i, m)
    st_n, st_t, _ = oracle.ref_build(1, v, i, m)
    np.savez_compressed(os.path.join(HERE, "boxes_bvh.npz"), stackless_nodes=sl_n, stackless_tris=sl_t,
                        stack_nodes=st_n, stack_tris=st_t)
    scene_defs = {
        "cloud12_s2": scenes.icosphere_cloud(count=12, subdivisions=2, seed=5),
        "cloud64_s3": scenes.icosphere_cloud(count=64, subdivisions=3, seed=9),
    }
    for name, meshes in scene_defs.items():
        v, i, m = concat_meshes(meshes)
        for fmt, key in ((0, "stackless"), (1, "stack")):
            nodes, tris, info = oracle.ref_build(fmt, v, i, m)
            out[f"{name}_{key}"] = {"scene": name, "format": key, "array_length": int(info["array_length"]),
                                    "nodes_sha256": sha(nodes), "tris_sha256": sha(tris)}
    with open(os.path.join(HERE, "bvh_hashes.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
