"""GPU parity of the batched persistent kernel: many scenes and ray types in one launch.

Every batch of a batched launch must produce the same hit records as the host
oracle and as a plain crt_trace of that batch alone, with and without hipGraph
replay, for every tuning configuration the autotuner may choose.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def work():
    import candela_amd as crt
    from candela_amd import workloads
    if crt.device_count() < 1:
        pytest.fail("no HIP device visible")
    log("[case] building 8 small batched scenes")
    return workloads.batched_scenes(crt, count=8, seed=5, tri_range=(1e3, 6e4), ray_range=(4e3, 3e4))


def oracle_hits(orc, s, which):
    hb = s.host
    if which == "closest":
        return orc.trace(0, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], 0, s.closest)
    return orc.trace(0, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], 1, s.shadow)


def check_all(work, orc, label):
    for s in work.scenes:
        for which, h in (("closest", s.h_closest), ("shadow", s.h_shadow)):
            got = h.download()
            ref = oracle_hits(orc, s, which)
            same = got.tobytes() == ref.tobytes()
            if not same:
                n = np.count_nonzero(got.view(np.uint8).reshape(-1, 32).any(1) != ref.view(np.uint8).reshape(-1, 32).any(1))
                diff = np.count_nonzero((got.view(np.uint32).reshape(-1, 8) != ref.view(np.uint32).reshape(-1, 8)).any(1))
                raise AssertionError(f"{label}: scene {s.index} {which}: {diff} hit records differ ({n})")
    log(f"[parity] {label}: {work.rays} rays over {len(work.scenes)} scenes bit-identical to the oracle")


def clear_hits(work):
    for s in work.scenes:
        s.h_closest.fill_bytes(0xAB)
        s.h_shadow.fill_bytes(0xAB)


def test_batched_default_plan(work, orc):
    import candela_amd as crt
    clear_hits(work)
    plan = crt.BatchPlan(work.batches(crt))
    assert plan.rays == work.rays
    plan.launch()
    crt.lib().crt_device_sync()
    check_all(work, orc, "batched default")


@pytest.mark.parametrize("variant", range(8))
def test_batched_configs(work, orc, variant):
    import candela_amd as crt
    for tpc, wpg, order, grid, shards in ((1, 4, 1, 0, 0), (2, 1, 0, 0, 16), (4, 2, 1, 3, 16), (1, 1, 1, 5, 16)):
        clear_hits(work)
        cfg = crt.BatchConfig(variant, tpc, wpg, grid, order, shards)
        plan = crt.BatchPlan(work.batches(crt), cfg)
        plan.launch()
        crt.lib().crt_device_sync()
        check_all(work, orc, f"batched variant={variant} tpc={tpc} wpg={wpg} order={order} grid={grid} "
                             f"shards={shards}")


def test_batched_graph_replays(work, orc):
    """Graph replays interleaved with eager plain traces and synchronisations; every replay is checked."""
    import candela_amd as crt
    plan = crt.BatchPlan(work.batches(crt))
    stream = crt.Stream()
    for rep in range(3):
        clear_hits(work)
        plan.launch(stream=stream, graph=True)
        stream.sync()
        check_all(work, orc, f"graph replay {rep}")
        # eager work between replays: plain traces of two batches
        s = work.scenes[rep % len(work.scenes)]
        crt.trace(s.scene, crt.QUERY_ANY, s.d_shadow, s.h_shadow, kernel=crt.KERNEL_WAVE)
        crt.lib().crt_device_sync()


def test_batched_equals_plain_with_counters(work):
    import candela_amd as crt
    s = work.scenes[-1]
    n = s.closest.size
    tt_b = crt.DeviceArray(np.uint32, n)
    tt_p = crt.DeviceArray(np.uint32, n)
    hb = crt.DeviceArray(crt.HIT_DTYPE, n)
    crt.BatchPlan([(s.scene, crt.QUERY_CLOSEST, s.d_closest, hb, 0, tt_b)]).launch()
    crt.trace(s.scene, crt.QUERY_CLOSEST, s.d_closest, s.h_closest, kernel=crt.KERNEL_THREAD_PER_RAY, tri_tests=tt_p)
    crt.lib().crt_device_sync()
    assert hb.download().tobytes() == s.h_closest.download().tobytes()
    assert np.array_equal(tt_b.download(), tt_p.download())


def test_autotune_returns_a_working_config(work, orc):
    import candela_amd as crt
    cands = [crt.BatchConfig(v, t, w, 0, o, q) for v in (0, 1) for t in (1, 2) for w in (1, 4) for o in (0, 1)
             for q in (1, 16)]
    best, ms = crt.autotune_batched(work.batches(crt), cands, repeats=1)
    log(f"[autotune] best variant={best.variant} tpc={best.tiles_per_claim} wpg={best.waves_per_group} "
        f"order={best.order} {ms:.3f} ms")
    assert ms > 0
    clear_hits(work)
    crt.BatchPlan(work.batches(crt), best).launch()
    crt.lib().crt_device_sync()
    check_all(work, orc, "autotuned config")


def wide_oracle_hits(orc, s, layout, which, child_order=0):
    """Host wide walk of one batch; child_order 1 walks the any-hit batch with CRT_FLAG_SLOT_ORDER."""
    import candela_amd as crt
    w = s.wide[layout].host
    rays, q = (s.closest, 0) if which == "closest" else (s.shadow, 1)
    flags = crt.FLAG_SLOT_ORDER if child_order == 1 and q == 1 else 0
    return orc.trace_wide(w.width, w.groups, w.records, w.entities, q, rays, flags=flags)


@pytest.mark.parametrize("width", [4, 8])
def test_batched_wide(work, orc, width):
    """The tile queue over the BVH4 / BVH8 layouts: every preset, both child orders, with and without graph
    replay."""
    import candela_amd as crt
    layout = (width, 2)
    work.build_wide(crt, width, 2)
    for variant, tpc, wpg, order, shards, co in ((0, 1, 4, 1, 16, 0), (1, 2, 1, 0, 1, 1), (2, 4, 2, 1, 16, 0),
                                                 (3, 1, 4, 0, 0, 1), (4, 2, 4, 1, 16, 1), (5, 1, 1, 1, 16, 0),
                                                 (6, 2, 4, 1, 16, 1), (7, 1, 2, 0, 16, 0)):
        for s in work.scenes:
            s.h_closest.fill_bytes(0xAB)
            s.h_shadow.fill_bytes(0xAB)
        plan = crt.BatchPlan(work.batches(crt, layout), crt.BatchConfig(variant, tpc, wpg, 0, order, shards, co))
        for replay in range(2):
            plan.launch(graph=replay == 1)
            crt.lib().crt_device_sync()
            for s in work.scenes:
                for which, h in (("closest", s.h_closest), ("shadow", s.h_shadow)):
                    got, ref = h.download(), wide_oracle_hits(orc, s, layout, which, co)
                    assert got.tobytes() == ref.tobytes(), f"wide{width} v{variant} co{co} scene {s.index} {which}"
        log(f"[parity] batched wide{width} variant={variant} tpc={tpc} wpg={wpg} order={order} shards={shards} "
            f"child_order={co}: {work.rays} rays bit-identical to the host wide walk (eager and graph)")
    # mixing node formats in one plan is rejected, and so is an unknown child order
    mixed = work.batches(crt, layout)[:2] + work.batches(crt)[2:4]
    with pytest.raises(crt.CrtError):
        crt.BatchPlan(mixed)
    with pytest.raises(crt.CrtError):
        crt.BatchPlan(work.batches(crt, layout), crt.BatchConfig(0, 1, 4, 0, 1, 16, 2))
    with pytest.raises(crt.CrtError):
        crt.BatchPlan(work.batches(crt, layout), crt.BatchConfig(12, 1, 4, 0, 1, 16, 0))


def span_twin(s, which, fn):
    """The host twin `fn(query, rays)` replayed per batch, as the plan's tiles run: the closest-hit rays of a scene
    are two batches (primary, then GI) whose tiles start at each batch's first ray."""
    if which == "shadow":
        return fn(1, s.shadow)
    parts = [s.closest[:s.primary], s.closest[s.primary:s.primary + s.gi]]
    return np.concatenate([fn(0, p) for p in parts if p.size])


# batched speculative presets -> crt_trace's CRT_KERNEL_SPECULATIVE presets (the host twin's schedules)
SPEC_OF_BATCHED = {8: 0, 9: 1, 10: 3, 11: 4}


@pytest.mark.parametrize("width", [4, 8])
def test_batched_wide_speculative(work, orc, width):
    """Presets 8..11 run the speculative walk per 64-ray tile: each tile is one wave of crt_trace's speculative
    grid launch on its batch, so every record, counters included, equals the host twin that replays that
    schedule -- eager and graph, for several claim sizes, group sizes, tile orders and shard counts."""
    import candela_amd as crt
    layout = (width, 2)
    work.build_wide(crt, width, 2)
    ref = {}
    for variant, tpc, wpg, order, shards, co in ((8, 1, 4, 1, 16, 0), (9, 2, 1, 0, 1, 1), (10, 4, 2, 1, 16, 0),
                                                 (11, 1, 4, 0, 16, 1)):
        for s in work.scenes:
            s.h_closest.fill_bytes(0xAB)
            s.h_shadow.fill_bytes(0xAB)
        plan = crt.BatchPlan(work.batches(crt, layout), crt.BatchConfig(variant, tpc, wpg, 0, order, shards, co))
        for replay in range(2):
            plan.launch(graph=replay == 1)
            crt.lib().crt_device_sync()
            for s in work.scenes:
                w = s.wide[layout].host
                for which, h in (("closest", s.h_closest), ("shadow", s.h_shadow)):
                    key = (s.index, which, variant)
                    if key not in ref:
                        ref[key] = span_twin(s, which, lambda q_, r_: orc.trace_wide_spec(
                            w.width, w.groups, w.records, w.entities, q_, r_, variant=SPEC_OF_BATCHED[variant]))
                    got = h.download()
                    assert got.tobytes() == ref[key].tobytes(), \
                        f"speculative wide{width} v{variant} scene {s.index} {which} (replay {replay})"
        log(f"[parity] batched speculative wide{width} variant={variant} tpc={tpc} wpg={wpg} order={order} "
            f"shards={shards}: {work.rays} rays bit-identical to the twin of the speculative grid launch "
            "(eager and graph)")


def test_batched_stackless_speculative(work, orc):
    """Presets 8..11 over the stackless layout run its speculative walk per 64-ray tile: one wave of crt_trace's
    speculative grid launch (presets 0..3), so every record, counters included, equals the host twin of that
    launch -- eager and graph, for several claim sizes, group sizes, tile orders and shard counts."""
    import candela_amd as crt
    ref = {}
    for variant, tpc, wpg, order, shards in ((8, 1, 4, 1, 16), (9, 2, 1, 0, 1), (10, 4, 2, 1, 16), (11, 1, 4, 0, 16)):
        clear_hits(work)
        plan = crt.BatchPlan(work.batches(crt), crt.BatchConfig(variant, tpc, wpg, 0, order, shards, 0))
        for replay in range(2):
            plan.launch(graph=replay == 1)
            crt.lib().crt_device_sync()
            for s in work.scenes:
                hb = s.host
                for which, h in (("closest", s.h_closest), ("shadow", s.h_shadow)):
                    key = (s.index, which, variant)
                    if key not in ref:
                        ref[key] = span_twin(s, which, lambda q_, r_: orc.trace_stackless_spec(
                            hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], q_, r_, variant=variant - 8))
                    assert h.download().tobytes() == ref[key].tobytes(), \
                        f"speculative stackless v{variant} scene {s.index} {which} (replay {replay})"
        log(f"[parity] batched speculative stackless variant={variant} tpc={tpc} wpg={wpg} order={order} "
            f"shards={shards}: {work.rays} rays bit-identical to the twin (eager and graph)")


def test_autotune_wide_child_order(work, orc):
    """The built-in candidate grid over a wide layout with any-hit batches includes both child orders; the
    returned config reproduces the host walk with its child order."""
    import candela_amd as crt
    layout = (4, 2)
    work.build_wide(crt, 4, 2)
    cands = [crt.BatchConfig(v, 2, 4, 0, 1, 16, co) for v in (0, 2) for co in (0, 1)]
    best, ms = crt.autotune_batched(work.batches(crt, layout), cands, repeats=1)
    log(f"[autotune] wide4 best variant={best.variant} child_order={best.child_order} {ms:.3f} ms")
    assert ms > 0 and best.child_order in (0, 1)
    for s in work.scenes:
        s.h_closest.fill_bytes(0xAB)
        s.h_shadow.fill_bytes(0xAB)
    crt.BatchPlan(work.batches(crt, layout), best).launch()
    crt.lib().crt_device_sync()
    for s in work.scenes:
        for which, h in (("closest", s.h_closest), ("shadow", s.h_shadow)):
            assert h.download().tobytes() == wide_oracle_hits(orc, s, layout, which, best.child_order).tobytes()


def test_batched_stack_format(orc):
    """Batches over stack-layout scenes (FlattenStackBVH): every preset and workgroup size, eager and graph,
    bit-identical to crt_trace's one-ray-per-lane stack walk and to the host walk; counters too."""
    import candela_amd as crt
    from candela_amd import scenes
    from oracle_cases import random_rays
    log("[case] batched stack format")
    ctxs, batches, refs, bufs = [], [], [], []
    for k in range(5):
        r = crt.RayIntersector(crt.FORMAT_STACK)
        r.add_object(0, scenes.icosphere_cloud(count=20 + 30 * k, subdivisions=2, seed=40 + k, box=20.0))
        r.push_entity(0)
        r.buffer_entities()
        r.buffer_data()
        hb = r.host_buffers()
        ctxs.append(r)
        for q, n, tmax in ((crt.QUERY_CLOSEST, 3000 + 1700 * k, 1e6), (crt.QUERY_ANY, 900 + 500 * k, 3.0)):
            rays = random_rays(n, [-2, -2, -2], [22, 22, 22], seed=50 + 10 * k + q, tmax=tmax)
            dr = crt.DeviceArray.from_host(rays)
            dh = crt.DeviceArray(crt.HIT_DTYPE, n)
            dt = crt.DeviceArray(np.uint32, n)
            o, ot = orc.trace(1, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], q, rays,
                              with_tests=True)
            p = crt.DeviceArray(crt.HIT_DTYPE, n)
            crt.trace(r.scene(), q, dr, p, kernel=crt.KERNEL_THREAD_PER_RAY)
            crt.lib().crt_device_sync()
            assert p.download().tobytes() == o.tobytes()
            batches.append((r.scene(), q, dr, dh, 0, dt))
            refs.append((o, ot))
            bufs.append((dr, dh, dt, p))
    for variant in range(8):
        for wpg, tpc in ((1, 1), (4, 2), (2, 4)):
            for _, dh, dt, _ in bufs:
                dh.fill_bytes(0xAB)
                dt.fill_bytes(0xCD)
            plan = crt.BatchPlan(batches, crt.BatchConfig(variant, tpc, wpg, 0, 1, 16))
            for replay in range(2):
                plan.launch(graph=replay == 1)
                crt.lib().crt_device_sync()
                for (_, dh, dt, _), (o, ot) in zip(bufs, refs):
                    assert dh.download().tobytes() == o.tobytes(), f"stack v{variant} wpg{wpg}"
                    assert np.array_equal(dt.download(), ot)
    log(f"[parity] batched stack format: {sum(b[2].count for b in batches)} rays, 8 presets x 3 shapes, "
        "eager and graph, bit-identical")
    # presets 8..11: the speculative stack walk per tile, equal to the twin of crt_trace's speculative grid launch
    # (presets 0..3), counters included
    twins = []
    for k, (scene, q, dr, _, _, _) in enumerate(batches):
        hb = ctxs[k // 2].host_buffers()
        rays = dr.download()
        twins.append([orc.trace_stack_spec(hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], q, rays,
                                           variant=v, with_tests=True) for v in range(4)])
    for variant, wpg, tpc in ((8, 4, 1), (9, 1, 2), (10, 2, 4), (11, 4, 2)):
        for _, dh, dt, _ in bufs:
            dh.fill_bytes(0xAB)
            dt.fill_bytes(0xCD)
        plan = crt.BatchPlan(batches, crt.BatchConfig(variant, tpc, wpg, 0, 1, 16))
        for replay in range(2):
            plan.launch(graph=replay == 1)
            crt.lib().crt_device_sync()
            for (_, dh, dt, _), tw in zip(bufs, twins):
                o, ot = tw[variant - 8]
                assert dh.download().tobytes() == o.tobytes(), f"speculative stack v{variant} wpg{wpg}"
                assert np.array_equal(dt.download(), ot)
    log("[parity] batched speculative stack format: presets 8..11, eager and graph, bit-identical to the twin")


def test_autotune_picks_a_layout(work, orc):
    """crt_autotune_batched_layouts over the BVH2, BVH4 and BVH8 layouts of the same scenes: the chosen layout's
    plan gives that layout's host-oracle records."""
    import candela_amd as crt
    for key in ((4, 2), (8, 2)):
        work.build_wide(crt, *key)
    keys = [None, (4, 2), (8, 2)]
    layouts = [work.batches(crt, k) for k in keys]
    cands = [crt.BatchConfig(v, 1, 4, 0, 1, 16, co) for v in (0, 6) for co in (0, 1)]
    idx, best, ms = crt.autotune_batched_layouts(layouts, cands, repeats=1)
    log(f"[autotune] layouts: best {keys[idx]} variant={best.variant} child_order={best.child_order} {ms:.3f} ms")
    assert 0 <= idx < 3 and ms > 0
    clear_hits(work)
    crt.BatchPlan(layouts[idx], best).launch()
    crt.lib().crt_device_sync()
    key = keys[idx]
    for s in work.scenes:
        if key is None:
            hb = s.host
            rc = orc.trace(0, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], 0, s.closest)
            ra = orc.trace(0, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], 1, s.shadow)
        else:
            w = s.wide[key].host
            rc = orc.trace_wide(w.width, w.groups, w.records, w.entities, 0, s.closest)
            ra = orc.trace_wide(w.width, w.groups, w.records, w.entities, 1, s.shadow,
                                flags=4 if best.child_order == 1 else 0)
        assert s.h_closest.download().tobytes() == rc.tobytes(), f"scene {s.index} closest ({key})"
        assert s.h_shadow.download().tobytes() == ra.tobytes(), f"scene {s.index} shadow ({key})"
    log(f"[parity] autotuned layout {key}: every batch bit-identical to its oracle")
