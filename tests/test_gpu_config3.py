"""GPU parity on the batched heterogeneous workload at its full size.

64 icosphere clouds with triangle counts log-uniform in [1k, 2M] (seed 3),
each with its own SAH BVH; per scene a ray batch of log-uniform size in
[4k, 2M], 40 % coherent primary, 40 % cosine GI (closest hit) and 20 % shadow
rays toward a random point light (any hit).  All 192 batches run in one
batched persistent launch replayed from a hipGraph, over the BVH2 and over
the BVH4 layout; every record of every batch must equal the host oracle (the
reference walk over the BVH2, the host twin of the wide walk over the BVH4).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def work(crt):
    from candela_amd import workloads
    if crt.device_count() < 1:
        pytest.fail("no HIP device visible")
    log("[case] config 3: building 64 scenes")
    w = workloads.batched_scenes(crt, count=64, seed=3)
    log(f"[case] config 3: {len(w.scenes)} scenes, {w.triangles} triangles, {w.rays} rays")
    assert len(w.scenes) == 64 and w.rays > 10_000_000
    return w


def _differs(a, b):
    return int(np.count_nonzero((a.view(np.uint32).reshape(-1, 8) != b.view(np.uint32).reshape(-1, 8)).any(1)))


def _check(work, orc, layout, child_order, label):
    bad = []
    for s in work.scenes:
        if layout is None:
            hb = s.host
            rc = orc.trace(0, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], 0, s.closest)
            ra = orc.trace(0, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], 1, s.shadow)
        else:
            w = s.wide[layout].host
            rc = orc.trace_wide(w.width, w.groups, w.records, w.entities, 0, s.closest)
            ra = orc.trace_wide(w.width, w.groups, w.records, w.entities, 1, s.shadow,
                                flags=4 if child_order == 1 else 0)
        dc, da = _differs(s.h_closest.download(), rc), _differs(s.h_shadow.download(), ra)
        if dc or da:
            bad.append((s.index, dc, da))
    assert not bad, f"{label}: (scene, closest, shadow) records differing: {bad[:8]}"
    log(f"[parity] {label}: {work.rays} rays over {len(work.scenes)} scenes bit-identical to the oracle")


@pytest.mark.parametrize("layout", [None, (4, 2)], ids=["bvh2", "bvh4"])
def test_config3_one_graph_replay(crt, orc, work, layout):
    if layout is not None:
        log(f"[case] config 3: wide layouts {layout} built in {work.build_wide(crt, *layout):.1f}s")
    # the configurations the autotuner picked for this workload: occupancy preset 6, one tile per claim,
    # waves per group 4, heaviest first, 16 queue shards (any-hit children in slot order over BVH4)
    child_order = 1 if layout is not None else 0
    cfg = crt.BatchConfig(6, 1, 4, 0, 1, 16, child_order)
    plan = crt.BatchPlan(work.batches(crt, layout), cfg)
    stream = crt.Stream()
    for s in work.scenes:
        s.h_closest.fill_bytes(0xAB)
        s.h_shadow.fill_bytes(0xAB)
    crt.lib().crt_device_sync()
    plan.launch(stream=stream, graph=True)
    stream.sync()
    _check(work, orc, layout, child_order, f"config 3 {'bvh2' if layout is None else 'bvh4'} graph replay")
    stream.close()
