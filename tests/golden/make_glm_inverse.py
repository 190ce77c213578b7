"""Writes tests/golden/glm_inverse.npz: entity transforms and glm::inverse of each.

    python tests/golden/make_glm_inverse.py

Compiles glm_inverse.cpp with g++ against the reference's vendored glm
headers (Source/Dependencies/glm, 0.9.8.5) into oracle/_ref/glm_inverse, feeds
it seeded model matrices and stores inputs and outputs.  The outputs are what
PushEntity (Intersector.h:209) writes into BVHEntity.InverseMatrix for those
models; libcandela_rt is not used here.
"""
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
GLM = "/root/reference/Source/Dependencies/glm"
EXE = os.path.join(ROOT, "oracle", "_ref", "glm_inverse")


def rotation(axis, angle):
    a = np.asarray(axis, np.float64)
    a /= np.linalg.norm(a)
    x, y, z = a
    c, s = np.cos(angle), np.sin(angle)
    C = 1 - c
    return np.array([[c + x * x * C, x * y * C - z * s, x * z * C + y * s],
                     [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
                     [z * x * C - y * s, z * y * C + x * s, c + z * z * C]])


def models(seed=17):
    """Row-major (math) 4x4 models: identity, translations, rotations, scales, shears, compositions and dense
    well-conditioned matrices over several magnitudes."""
    rng = np.random.default_rng(seed)
    out = [np.eye(4)]
    for _ in range(40):   # translate * rotate * scale, the usual entity transform
        m = np.eye(4)
        m[:3, :3] = rotation(rng.normal(size=3), rng.uniform(-np.pi, np.pi)) @ np.diag(rng.uniform(0.05, 20, 3))
        m[:3, 3] = rng.uniform(-500, 500, 3)
        out.append(m)
    for _ in range(20):   # pure translations and uniform scales
        m = np.eye(4) * rng.uniform(0.01, 100)
        m[3, 3] = 1.0
        m[:3, 3] = rng.normal(scale=1e3, size=3)
        out.append(m)
    for _ in range(20):   # shears
        m = np.eye(4)
        m[:3, :3] += np.triu(rng.normal(scale=0.5, size=(3, 3)), 1)
        m[:3, 3] = rng.normal(size=3)
        out.append(m)
    for _ in range(40):   # dense, projective last row included
        while True:
            m = rng.normal(size=(4, 4)) * 10 ** rng.uniform(-2, 2)
            if np.linalg.cond(m) < 1e3:
                break
        out.append(m)
    # exactly representable entries (small integers and halves)
    for _ in range(20):
        m = np.eye(4)
        m[:3, :3] = rng.integers(-4, 5, size=(3, 3)) / 2.0
        if abs(np.linalg.det(m)) < 0.5:
            continue
        m[:3, 3] = rng.integers(-8, 9, size=3)
        out.append(m)
    return np.stack(out).astype(np.float32)


def main():
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", f"-I{GLM}",
                    os.path.join(HERE, "glm_inverse.cpp"), "-o", EXE], check=True)
    m = models()
    colmajor = np.ascontiguousarray(np.transpose(m, (0, 2, 1))).reshape(-1, 16)
    res = subprocess.run([EXE], input=colmajor.tobytes(), stdout=subprocess.PIPE, check=True)
    inv = np.frombuffer(res.stdout, np.float32).reshape(-1, 16)
    assert inv.shape == colmajor.shape
    np.savez_compressed(os.path.join(HERE, "glm_inverse.npz"), model=colmajor, inverse=inv)
    print(f"glm_inverse.npz: {len(m)} models")


if __name__ == "__main__":
    main()
