import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def concat_meshes(meshes):
    """(vertices, indices with vertex offsets, mesh id per triangle) as BuildBVH concatenates them."""
    vs, idx, mids, off = [], [], [], 0
    for v, i, m in meshes:
        vs.append(v)
        idx.append(np.asarray(i, np.uint32) + np.uint32(off))
        mids.append(np.full(len(i) // 3, m, np.int32))
        off += len(v)
    return np.concatenate(vs), np.concatenate(idx), np.concatenate(mids)


@pytest.fixture(scope="session")
def crt():
    import candela_amd
    candela_amd.lib()
    return candela_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not oracle.available():
        pytest.skip("oracle/liboracle.so not built (make -C oracle)")
    oracle.lib()
    return oracle
