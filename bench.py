#!/usr/bin/env python3
"""Incoherent-ray throughput of the traversal kernels on one MI355X.

Workload: 2,000 subdivision-3 icospheres (2.56M triangles, centres U[0,100]^3,
radii lognormal(0, 0.5), seed 2) in one SAH BVH built by the CPU constructor.
A 1920x1080 G-buffer comes from primary closest hits; every pixel that hit
spawns one cosine-weighted RTAO any-hit ray (tmax 2.0) and one cosine-weighted
GI closest-hit ray (tmax 1e6), origins offset by 1e-4 along the face normal,
directions from a per-pixel PCG hash.  A step traces both batches (all rays
resident in HBM before timing starts); value = rays traced / second.

The candidates are the BVH2 kernels (one ray per lane; wave walk and
persistent presets) and the one-ray-per-lane walk over the BVH4 / BVH8 layouts
derived from the same BVH2; value is the best.  The dominant kernel (GI
closest hit) is priced against HBM bandwidth with the algorithmic bytes of
every ray counted exactly by the counting build of the kernel: 32 B per
stackless node visit (64 B per stack node, 128 B per BVH4 node, 256 B per
BVH8 node), 64 B per triangle test (16-B triangle + three 16-B positions; 48
B per triangle record in the wide layouts), 32 B ray read + 32 B hit write.

cpu_baseline: the host restatement of the traversal (oracle/, the reference's
GLSL walk in C) timed on a bounded sample of the same rays.

    python bench.py [--steps K] [--warmup W] [--format stackless|stack]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "incoherent Mrays/s, batched vs plain launch on 1 MI355X; % of HBM roofline"
HBM_PEAK_GBS = 8000.0
KERNEL_NAMES = {1: "trace_kernel", 2: "persistent_stackless", 3: "grid_stackless", 4: "sbvh2::grid (speculative)",
                5: "sbvh2::persistent (speculative, lane refill)"}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Events:
    def __init__(self, crt, n):
        import ctypes as C
        self.crt = crt
        self.C = C
        self.ev = []
        for _ in range(n):
            h = C.c_void_p()
            crt._lib.check(crt.lib().crt_event_create(C.byref(h)))
            self.ev.append(h.value)

    def record(self, i, stream):
        self.crt._lib.check(self.crt.lib().crt_event_record(self.ev[i], stream))

    def ms(self, a, b) -> float:
        out = self.C.c_float()
        self.crt._lib.check(self.crt.lib().crt_event_elapsed_ms(self.ev[a], self.ev[b], self.C.byref(out)))
        return float(out.value)

    def close(self):
        for e in self.ev:
            self.crt.lib().crt_event_destroy(e)


def bytes_per_ray(fmt: int, visits: np.ndarray, tests: np.ndarray) -> np.ndarray:
    """Algorithmic bytes: node bytes per visit + triangle bytes per test + 32-B ray read + 32-B hit write.
    BVH2: 32-B stackless / 64-B stack node, 64 B per test (16-B triangle + three 16-B positions).
    Wide: 128 B per group of four children (one group for BVH4, two for BVH8), 48-B triangle record."""
    if fmt in (2, 3):
        node, tri = (128.0 if fmt == 2 else 256.0), 48.0
    else:
        node, tri = (32.0 if fmt == 0 else 64.0), 64.0
    return node * visits.astype(np.float64) + tri * tests.astype(np.float64) + 64.0


def _kernel_name(layout: str, kernel: int, variant: int) -> str:
    if layout == "stack":
        return {1: "trace_kernel (stack walk)", 3: "stack_wave, preset %d" % variant,
                2: "stack_persistent (lane refill), preset %d" % variant,
                4: "sbvh2::grid (speculative, stack layout), preset %d" % variant,
                5: "sbvh2::persistent (speculative, stack layout, lane refill), preset %d" % variant}.get(
                    kernel, "trace_kernel")
    if layout == "bvh2":
        return "%s, preset %d" % (KERNEL_NAMES[kernel], variant)
    return {3: "wide_grid, preset %d" % variant, 2: "wide_persistent (lane refill), preset %d" % variant,
            4: "spec_grid (speculative, wave-spread leaf tests), preset %d" % variant,
            5: "spec_persistent (speculative, lane refill), preset %d" % variant}.get(kernel, "trace_kernel")


def kernel_label(c: dict, gi_only: bool = False) -> str:
    name = _kernel_name(c["layout"], c["kernel"], c["variant"])
    if c["layout"] == "bvh2":
        label = "plain (one batch per launch): %s, BVH2 stackless" % name
    elif c["layout"] == "stack":
        label = "plain (one batch per launch): %s, BVH2 stack layout (FlattenStackBVH), LDS stack %d" % (
            name, c.get("lds", 16))
    else:
        label = "plain (one batch per launch): %s over the BVH%d layout derived from the BVH2 (%s collapse), " \
                "leaves of up to %d triangles" % (name, c["width"], "SAH" if c.get("collapse") else "greedy",
                                                  c["max_leaf"])
    if "ao_kernel" in c and not gi_only:
        label += "; AO launch: " + _kernel_name(c["layout"], c["ao_kernel"], c["ao_variant"])
    if c.get("ao_order") == 1 and not gi_only:
        label += "; AO children in slot order"
    if c.get("streams") == 2 and not gi_only:
        label += "; AO and GI launches overlapped on two streams"
    return label


def choice_name(c: dict) -> str:
    """'K:V' for a BVH2 kernel choice, 'wWlL:K:V' for kernel K, variant V over the wide layout of width W
    and leaf size L ('wWsL' with the SAH collapse); a choice whose AO launch uses another kernel gets '+ao:K:V', and ':s' after it when the
    AO walks take children in slot order (CRT_FLAG_SLOT_ORDER); '+overlap' when the AO and GI launches run on two
    streams."""
    if c["layout"] == "bvh2":
        name = f"{c['kernel']}:{c['variant']}"
    elif c["layout"] == "stack":
        name = f"s{c.get('lds', 16)}:{c['kernel']}:{c['variant']}"
    else:
        name = f"w{c['width']}{'s' if c.get('collapse') else 'l'}{c['max_leaf']}:{c['kernel']}:{c['variant']}"
    if "ao_kernel" in c:
        name += f"+ao:{c['ao_kernel']}:{c['ao_variant']}"
    if c.get("ao_order") == 1:
        name += ":s" if "ao_kernel" in c else f"+ao:{c['kernel']}:{c['variant']}:s"
    if c.get("streams") == 2:
        name += "+overlap"
    return name


# Cost of one triangle test against one wide-node visit for the SAH-collapsed layouts (crt_wide_build_ex).
SAH_TRI_COST = 0.3


def wide_key(c: dict) -> tuple:
    """(width, max_leaf, collapse) of a wide-layout choice."""
    return c["width"], c["max_leaf"], c.get("collapse", 0)


def layout_name(layout) -> str:
    """bvh2, wide4, wide8, or wide4_sah for the SAH collapse, of a batched-workload layout key."""
    if layout is None:
        return "bvh2"
    return f"wide{layout[0]}" + ("_sah" if len(layout) > 2 and layout[2] else "")


def gi_choice_name(c: dict) -> str:
    """The choice of the GI launch alone."""
    return choice_name({k: v for k, v in c.items() if not k.startswith("ao_") and k != "streams"})


def ao_of(c: dict) -> tuple:
    """(kernel, variant) of the AO launch."""
    return c.get("ao_kernel", c["kernel"]), c.get("ao_variant", c["variant"])


def ao_flags(crt, c: dict) -> int:
    """Trace flags of the AO launch: the child order of its any-hit walks."""
    return crt.FLAG_SLOT_ORDER if c.get("ao_order") == 1 else 0


def measured_traffic(best: dict, args, fmt: int):
    """HBM bytes per GI launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    (profiles/traffic_*.json) when they were taken on this kernel choice and workload."""
    import glob
    choice = gi_choice_name(best)
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "traffic_*.json"))):
        with open(path) as f:
            d = json.load(f)
        if d.get("kernel_choice") == choice and d.get("workload") == "cloud2000_s3_gi" and args.spheres == 2000 \
                and fmt == 0:
            extra = {k: d[k] for k in ("l2_hit_rate", "valu_issue_frac") if k in d}
            return float(d["traffic_bytes"]), os.path.relpath(path, ROOT), extra
    return None, None, {}

def time_batched(crt, batches, steps: int, warmup: int, stream, wide: bool) -> dict:
    """Autotunes the batched kernel on `batches`, then times `steps` hipGraph replays of the best plan."""
    best, tune_ms = crt.autotune_batched(batches, default_candidates(crt, wide), repeats=3, stream=stream)
    plan = crt.BatchPlan(batches, best)
    for _ in range(max(1, warmup)):
        plan.launch(stream=stream, graph=True)
    stream.sync()
    ev = Events(crt, 2)
    ev.record(0, stream.handle)
    for _ in range(steps):
        plan.launch(stream=stream, graph=True)
    ev.record(1, stream.handle)
    ms = ev.ms(0, 1) / steps
    ev.close()
    cfg = {k: getattr(best, k) for k in ("variant", "tiles_per_claim", "waves_per_group", "order", "shards",
                                         "child_order")}
    return {"ms": ms, "config": cfg, "tune_ms": tune_ms, "rays": plan.rays}


# batched presets 8..11 (the speculative walk per tile) -> crt_trace's CRT_KERNEL_SPECULATIVE presets over the wide
# layouts (over the BVH2 layouts: presets 0..3)
BATCHED_SPEC_PRESET = {8: 0, 9: 1, 10: 3, 11: 4}


def default_candidates(crt, wide: bool):
    """Autotuner grid: preset (8..11: the speculative walk), tiles per claim, waves per group, tile order, queue
    shards and, over the wide layouts, the child order of the sequential any-hit walks."""
    presets = (0, 1, 4, 6, 7, 8, 9, 10, 11)
    return [crt.BatchConfig(v, t, w, 0, o, q, co) for v in presets for t in (1, 2) for w in (1, 4)
            for o in (0, 1) for q in (1, 16) for co in ((0, 1) if wide and v < 8 else (0,))]


def batched_vs_plain(crt, args, plain_choice: dict, stream) -> dict:
    """64 scenes (log-uniform 1k..2M triangles) x (primary, GI, shadow) batches: plain launches vs one launch,
    over the BVH2 and over the BVH4 / BVH8 layouts derived from it (plain_choice: layout key -> (kernel, variant
    of the closest-hit launches, kernel, variant and child order of the any-hit launches))."""
    from candela_amd import workloads
    t0 = time.perf_counter()
    work = workloads.batched_scenes(crt, count=args.batched_scenes, seed=3)
    log(f"batched workload: {len(work.scenes)} scenes, {work.triangles} triangles, {work.rays} rays "
        f"(BVH builds {work.build_s:.1f}s, setup {time.perf_counter() - t0:.1f}s)")
    K = args.batched_steps
    per_layout = {}
    for layout, (kernel, variant, ao_kernel, ao_variant, ao_order) in plain_choice.items():
        if layout is not None:
            log(f"wide layouts {layout} of all scenes built in {work.build_wide(crt, *layout):.1f}s")
        kw = dict(kernel=kernel, variant=variant, stream=stream)
        kw_any = dict(kernel=ao_kernel, variant=ao_variant, stream=stream,
                      flags=crt.FLAG_SLOT_ORDER if ao_order == 1 else 0)

        def plain_step():
            for s in work.scenes:   # one launch per (scene, query): primary+GI closest, shadow any
                sc = s.scene_for(layout)
                crt.trace(sc, crt.QUERY_CLOSEST, s.d_closest, s.h_closest, **kw)
                crt.trace(sc, crt.QUERY_ANY, s.d_shadow, s.h_shadow, **kw_any)

        for _ in range(max(1, args.warmup)):
            plain_step()
        stream.sync()
        ev = Events(crt, 2)
        ev.record(0, stream.handle)
        for _ in range(K):
            plain_step()
        ev.record(1, stream.handle)
        plain_ms = ev.ms(0, 1) / K
        ev.close()
        timed_hits = [[s.h_closest.download(), s.h_shadow.download()] for s in work.scenes]
        b = time_batched(crt, work.batches(crt, layout), K, args.warmup, stream, wide=layout is not None)
        batched_hits = [(s.h_closest.download(), s.h_shadow.download()) for s in work.scenes]
        # the batched records byte for byte against plain launches of the walk the chosen plan runs: the
        # speculative grid launch with the matching preset (a tile is one of its waves), else the wave walk
        v_b = b["config"]["variant"]
        spec_preset = BATCHED_SPEC_PRESET.get(v_b) if layout is not None else v_b - 8
        ref_kw = dict(kernel=crt.KERNEL_SPECULATIVE, variant=spec_preset) if v_b >= 8 else \
            dict(kernel=crt.KERNEL_WAVE, variant=0)
        ref_flags = crt.FLAG_SLOT_ORDER if (layout is not None and b["config"]["child_order"] == 1) else 0
        for s in work.scenes:
            sc = s.scene_for(layout)
            crt.trace(sc, crt.QUERY_CLOSEST, s.d_closest, s.h_closest, stream=stream, **ref_kw)
            crt.trace(sc, crt.QUERY_ANY, s.d_shadow, s.h_shadow, stream=stream, flags=ref_flags, **ref_kw)
        stream.sync()
        plain_hits = [[s.h_closest.download(), s.h_shadow.download()] for s in work.scenes]
        # the timed plain kernel against the batched records: nearest t and occlusion
        equivalent = all(np.array_equal(bh[0]["t"], th[0]["t"]) and
                         np.array_equal(bh[1]["t"] > 0, th[1]["t"] > 0)
                         for bh, th in zip(batched_hits, timed_hits))
        same = all(bh[0].tobytes() == ph[0].tobytes() and bh[1].tobytes() == ph[1].tobytes()
                   for bh, ph in zip(batched_hits, plain_hits))
        name = layout_name(layout)
        per_layout[name] = {"plain_ms": round(plain_ms, 4), "plain_mrays_s": round(work.rays / plain_ms / 1e3, 2),
                            "plain_kernel": f"{kernel}:{variant}" + (
                                f"+any:{ao_kernel}:{ao_variant}" + (":s" if ao_order == 1 else "")
                                if (ao_kernel, ao_variant, ao_order) != (kernel, variant, 0) else ""),
                            "batched_ms": round(b["ms"], 4),
                            "batched_mrays_s": round(work.rays / b["ms"] / 1e3, 2),
                            "speedup": round(plain_ms / b["ms"], 3), "batched_config": b["config"],
                            "hits_identical": bool(same), "timed_plain_t_and_occlusion_equal": bool(equivalent)}
        log(f"{name}: plain {2 * len(work.scenes)} launches {plain_ms:.3f} ms ({work.rays / plain_ms / 1e3:.1f} "
            f"Mrays/s), batched one launch {b['ms']:.3f} ms ({work.rays / b['ms'] / 1e3:.1f} Mrays/s) -> "
            f"{plain_ms / b['ms']:.2f}x; identical hits: {same}")
    # the library's own node-width choice: crt_autotune_batched_layouts over every layout built above
    keys = list(plain_choice.keys())
    t0 = time.perf_counter()
    idx, cfg, ms = crt.autotune_batched_layouts([work.batches(crt, k) for k in keys], default_candidates(crt, True),
                                                repeats=2, stream=stream)
    lib_choice = {"layout": layout_name(keys[idx]), "ms": round(ms, 4),
                  "config": {k: getattr(cfg, k) for k in ("variant", "tiles_per_claim", "waves_per_group", "order",
                                                          "shards", "child_order")},
                  "tune_s": round(time.perf_counter() - t0, 1)}
    log(f"library layout autotuner: {lib_choice}")
    best_plain = min(per_layout.items(), key=lambda kv: kv[1]["plain_ms"])
    best_batched = min(per_layout.items(), key=lambda kv: kv[1]["batched_ms"])
    return {"workload": f"{len(work.scenes)} icosphere clouds, {work.triangles} triangles, {work.rays} rays "
                        "(40% primary, 40% GI closest, 20% shadow any-hit)",
            "plain_launches": 2 * len(work.scenes),
            "plain_ms": best_plain[1]["plain_ms"], "plain_layout": best_plain[0],
            "plain_mrays_s": best_plain[1]["plain_mrays_s"],
            "batched_ms": best_batched[1]["batched_ms"], "batched_layout": best_batched[0],
            "batched_mrays_s": best_batched[1]["batched_mrays_s"],
            "speedup": round(best_plain[1]["plain_ms"] / best_batched[1]["batched_ms"], 3),
            "speedup_same_layout": {k: v["speedup"] for k, v in per_layout.items()},
            "hits_identical": all(v["hits_identical"] for v in per_layout.values()),
            "timed_plain_t_and_occlusion_equal": all(v["timed_plain_t_and_occlusion_equal"]
                                                     for v in per_layout.values()),
            "library_autotuner": lib_choice,
            "layouts": per_layout}


def time_attributes(crt, scene, d_hits, n: int, hits: np.ndarray, stream, steps: int) -> dict:
    """crt_hit_attributes over the GI hits: every sphere (mesh) gets a model colour; every other one an
    albedo texture (8 textures of 256x256 RGBA8).  Algorithmic bytes per hit: 32 (hit) + 32 (written
    record), and for a hit: 16 (triangle) + 3 x 16 (packed vertex halves) + 32 (material) + 8 (entity
    data), + 16 (texture descriptor) + 4 x 4 (texels) when textured."""
    rng = np.random.default_rng(7)
    nmesh = int(hits["mesh"].max()) + 1 if (hits["t"] > 0).any() else 1
    refs = np.zeros(nmesh, crt.TEXTURE_REF_DTYPE)
    refs["model_color"] = rng.uniform(0, 1, size=(nmesh, 4)).astype(np.float32)
    refs["albedo"] = np.where(np.arange(nmesh) % 2 == 0, np.arange(nmesh) % 8, -1)
    refs["normal"] = -1
    textures = [rng.integers(0, 2 ** 32, size=(256, 256), dtype=np.uint64).astype(np.uint32) for _ in range(8)]
    mats = crt.MaterialTable(refs, textures)
    out = crt.DeviceArray(crt.HIT_ATTR_DTYPE, n)
    crt.hit_attributes(scene, mats, d_hits, out, n, stream=stream)
    stream.sync()
    ev = Events(crt, 2)
    ev.record(0, stream.handle)
    for _ in range(steps):
        crt.hit_attributes(scene, mats, d_hits, out, n, stream=stream)
    ev.record(1, stream.handle)
    ms = ev.ms(0, 1) / steps
    ev.close()
    hit = hits["t"] > 0
    textured = hit & (refs["albedo"][np.clip(hits["mesh"], 0, nmesh - 1)] >= 0)
    nbytes = 64.0 * n + 136.0 * hit.sum() + 32.0 * textured.sum()
    res = {"hits": int(n), "hit_fraction": round(float(hit.mean()), 4), "textured_fraction": round(float(textured.mean()), 4),
           "ms": round(ms, 4), "mhits_s": round(n / ms / 1e3, 1), "bytes_per_launch": nbytes,
           "achieved_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1)}
    log(f"hit attributes: {n} GI hits in {ms:.4f} ms -> {res['mhits_s']} Mhits/s, {res['achieved_GBs']} GB/s algorithmic")
    return res


def time_shading(crt, stream, steps: int, size: int = 1024) -> dict:
    """Config 4: the differentiable Cook-Torrance + one-bounce shading of the boxes scene at size x size with 8
    point lights.  The forward pass traces, per pixel with a primary hit, one closest-hit bounce ray and 8 any-hit
    shadow rays from the hit, plus 8 from the bounce hit when there is one; the backward pass re-traces the 16
    shadow rays (visibility recomputed, not stored) and adds the gradients.  HIP events around `steps` launches
    of each pass on one stream."""
    from candela_amd import workloads
    pb = workloads.shading_problem(crt, size=size)
    n = size * size
    color = crt.DeviceArray(np.float32, 3 * n)
    target = crt.DeviceArray(np.float32, 3 * n)
    rec = crt.DeviceArray(crt.SHADE_RECORD_DTYPE, n)
    pb.model.set(pb.target_albedo, pb.target_roughness, pb.target_intensity)
    pb.model.forward(pb.scene, pb.rays, pb.hits, target, rec, stream=stream)
    pb.model.set(pb.albedo, pb.roughness, pb.intensity)
    ntri = pb.roughness.size
    ga = crt.DeviceArray(np.float32, 3 * ntri)
    gr = crt.DeviceArray(np.float32, ntri)
    gi = crt.DeviceArray(np.float32, pb.intensity.size)
    for d in (ga, gr, gi):
        d.fill_bytes(0)
    pb.model.forward(pb.scene, pb.rays, pb.hits, color, rec, stream=stream)
    pb.model.backward(pb.scene, rec, color, target, ga, gr, gi, stream=stream)
    stream.sync()
    r = rec.download()
    n_p = int(np.count_nonzero(r["tri0"] >= 0))
    n_q = int(np.count_nonzero(r["tri1"] >= 0))
    nl = int(pb.intensity.size)
    fwd_rays = n_p + nl * (n_p + n_q)
    bwd_rays = nl * (n_p + n_q)
    ev = Events(crt, 3)
    ev.record(0, stream.handle)
    for _ in range(steps):
        pb.model.forward(pb.scene, pb.rays, pb.hits, color, rec, stream=stream)
    ev.record(1, stream.handle)
    for _ in range(steps):
        pb.model.backward(pb.scene, rec, color, target, ga, gr, gi, stream=stream)
    ev.record(2, stream.handle)
    stream.sync()
    f_ms, b_ms = ev.ms(0, 1) / steps, ev.ms(1, 2) / steps
    ev.close()
    res = {"pixels": n, "lights": nl, "triangles": ntri, "forward_ms": round(f_ms, 4), "backward_ms": round(b_ms, 4),
           "step_ms": round(f_ms + b_ms, 4), "forward_rays": fwd_rays, "backward_rays": bwd_rays,
           "forward_mrays_s": round(fwd_rays / f_ms / 1e3, 1), "backward_mrays_s": round(bwd_rays / b_ms / 1e3, 1)}
    log(f"config 4 shading: forward {f_ms:.3f} ms ({res['forward_mrays_s']} Mrays/s), backward {b_ms:.3f} ms "
        f"({res['backward_mrays_s']} Mrays/s) at {size}x{size}, {nl} lights")
    return res


def time_raygen(crt, scene, d_rays, d_hits, rays: np.ndarray, hits: np.ndarray, hb: dict, stream, steps: int) -> dict:
    """crt_gen_diffuse_rays over the 1920x1080 primary G-buffer: AO + GI rays of every pixel that hit, compacted
    (count, scan, emit).  Algorithmic bytes: 32 (hit, count pass) + 64 (ray + hit, emit pass) per record, and
    per hit 16 (triangle) + 3 x 16 (positions) + 2 x 32 (two rays written).  The rays are checked against the
    host twin."""
    n = rays.size
    gen = crt.RayGen(n)
    ao, gi = crt.DeviceArray(crt.RAY_DTYPE, n), crt.DeviceArray(crt.RAY_DTYPE, n)
    p = crt.raygen_params(seed=2, offset=1e-4)
    gen.diffuse(scene, d_rays, d_hits, p, ao, gi, stream=stream)
    stream.sync()
    ev = Events(crt, 2)
    ev.record(0, stream.handle)
    for _ in range(steps):
        gen.diffuse(scene, d_rays, d_hits, p, ao, gi, stream=stream)
    ev.record(1, stream.handle)
    ms = ev.ms(0, 1) / steps
    ev.close()
    k = gen.last_count(stream)
    nbytes = 96.0 * n + 128.0 * k
    exact = None
    try:
        import oracle
        oao, ogi = oracle.gen_diffuse_rays(p, rays, hits, hb["entities"], tris=hb["triangles"], verts=hb["vertices"])
        exact = bool(k == oao.size and ao.download()[:k].tobytes() == oao.tobytes()
                     and gi.download()[:k].tobytes() == ogi.tobytes())
    except ImportError:
        pass
    res = {"records": int(n), "rays": 2 * k, "ms": round(ms, 4), "mrays_s": round(2 * k / ms / 1e3, 1),
           "bytes_per_launch": nbytes, "achieved_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1), "exact": exact}
    log(f"raygen: {2 * k} AO + GI rays from {n} primary hits in {ms:.4f} ms -> {res['mrays_s']} Mrays/s, "
        f"{res['achieved_GBs']} GB/s algorithmic, exact {exact}")
    return res


def time_gpu_build(crt, meshes, stream, steps: int) -> dict:
    """crt_build_bvh_gpu over one object (linear BVH over Morton codes): HIP-event time per build, after one
    warm-up build, with inputs resident on the device."""
    from candela_amd import scenes
    verts, idx, mids = scenes.concat_meshes(meshes)
    b = crt.GpuBVH(verts, idx, mids, stream=stream)
    stream.sync()
    ev = Events(crt, 2)
    ev.record(0, stream.handle)
    for _ in range(steps):
        b.build(stream)
    ev.record(1, stream.handle)
    ms = ev.ms(0, 1) / steps
    ev.close()
    log(f"gpu build: {b.ntris} triangles in {ms:.3f} ms -> {b.ntris / ms / 1e3:.1f} Mtris/s")
    return {"triangles": int(b.ntris), "ms": round(ms, 3), "mtris_s": round(b.ntris / ms / 1e3, 1)}, b


def time_large(crt, args) -> dict:
    """The large-scene workload: a 2048^2 fBm terrain (8,388,608 triangles) plus 1,400 subdivision-4
    icospheres (7,168,000), 3840 x 2160 pixels x 4 samples, 2 diffuse bounces, roulette p = 0.8 from the
    second bounce, sky light; one persistent launch traces every path (crt_trace_paths over the BVH4
    layout).  Mrays/s counts every traced segment.  2,000 sampled paths are checked against the host twin."""
    from candela_amd import scenes
    d = max(1, args.large_scale)
    t0 = time.perf_counter()
    meshes = scenes.large_scene(n=2048 // d, spheres=1400 // d)
    t1 = time.perf_counter()
    r = crt.RayIntersector(crt.FORMAT_STACKLESS)
    r.add_object(0, meshes)
    t2 = time.perf_counter()
    gpu_build = None
    try:
        gpu_build, gb = time_gpu_build(crt, meshes, crt.Stream(), 5)
        del gb
    except Exception as e:   # reported, never fatal to the workload
        gpu_build = {"error": f"{type(e).__name__}: {e}"}
    del meshes
    t2b = time.perf_counter()
    r.push_entity(0)
    r.buffer_entities()                              # materialises the entity record
    hb = r.host_buffers()
    lay = args.large_layout
    if len(lay) < 2 or lay[0] not in "ls" or not lay[1:].isdigit():
        raise SystemExit(f"--large-layout {lay}: expected lN or sN")
    sah = lay[0] == "s"
    # one host collapse, reused by the parity sample
    dw = crt.upload_wide(crt.FORMAT_STACKLESS, hb, 4, int(lay[1:]), collapse=crt.COLLAPSE_SAH if sah else 0,
                         tri_cost=SAH_TRI_COST if sah else 0.0)
    t3 = time.perf_counter()
    ntri = int(hb["triangles"].size)
    nmesh = int(hb["triangles"]["packed"][:, 3].max()) + 1
    rng = np.random.default_rng(5)
    albedo = rng.uniform(0.2, 0.8, size=(nmesh, 3)).astype(np.float32)
    albedo[0] = (0.45, 0.4, 0.3)                     # terrain
    dalb = crt.DeviceArray.from_host(albedo.ravel())
    W, H, spp = 3840 // d, 2160 // d, 4
    iv, ip = scenes.camera_inverses(aspect=W / H, **scenes.LARGE_CAMERA)
    p = crt.path_params(W, H, spp, iv, ip, albedo=dalb, seed=5, max_bounces=2, rr=0.8, sky=(1.0, 1.0, 1.0))
    n = W * H * spp
    out = crt.DeviceArray(crt.PATH_SAMPLE_DTYPE, n)
    stream = crt.Stream()
    scene = dw.scene
    best = None
    for v in (range(10) if args.large_preset < 0 else [args.large_preset]):
        crt.trace_paths(scene, p, out, n, variant=v, stream=stream)   # warm-up
        stream.sync()
        ev = Events(crt, 2)
        ev.record(0, stream.handle)
        reps = 3
        for _ in range(reps):
            crt.trace_paths(scene, p, out, n, variant=v, stream=stream)
        ev.record(1, stream.handle)
        ms = ev.ms(0, 1) / reps
        ev.close()
        log(f"large scene: path preset {v}: {ms:.3f} ms per frame")
        if best is None or ms < best[1]:
            best = (v, ms)
    crt.trace_paths(scene, p, out, n, variant=best[0], stream=stream)
    stream.sync()
    res = out.download()
    segs = int(res["segments"].sum())
    visits = res["visits"].astype(np.float64).sum()
    tests = res["tests"].astype(np.float64).sum()
    nbytes = 128.0 * visits + 48.0 * tests + 32.0 * n
    parity = None
    try:
        import oracle
        w = dw.host
        idx = np.random.default_rng(6).choice(n, 2000, replace=False)
        ref = oracle.trace_paths(p, 4, w.groups, w.records, w.entities, albedo, idx)
        if best[0] >= 6:
            # speculative presets: radiance, segments and status are the twin's; visit and test counts follow
            # the wave schedule
            got = res[idx]
            parity = {"paths": int(idx.size), "exact": None,
                      "radiance_segments_status_exact": bool(
                          got["radiance"].tobytes() == ref["radiance"].tobytes() and
                          np.array_equal(got["segments"], ref["segments"]) and
                          np.array_equal(got["status"], ref["status"]))}
        else:
            parity = {"paths": int(idx.size), "exact": bool(res[idx].tobytes() == ref.tobytes())}
    except ImportError:
        parity = None
    ms = best[1]
    out_d = {"workload": f"terrain {2048 // d}^2 + {1400 // d} icospheres ({ntri} triangles), {W}x{H}x{spp} paths, "
                         f"2 diffuse bounces, roulette 0.8, BVH4 layout {lay} ({dw.host.stats['node_count']} nodes)",
             "triangles": ntri, "paths": n, "segments": segs, "ms_per_frame": round(ms, 3),
             "mrays_s": round(segs / ms / 1e3, 1), "mpaths_s": round(n / ms / 1e3, 1), "preset": best[0],
             "mean_segments": round(segs / n, 4), "bytes_per_frame": nbytes,
             "achieved_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
             "status_flags": int(np.count_nonzero(res["status"])),
             "setup_s": {"scene": round(t1 - t0, 1), "bvh": round(t2 - t1, 1), "upload_and_wide": round(t3 - t2b, 1)},
             "gpu_build": gpu_build,
             "parity": parity}
    log(f"large scene: {out_d['mrays_s']} Mrays/s ({segs} segments in {ms:.2f} ms), parity {parity}")
    return out_d


def host_cpu_info() -> dict:
    """The host's logical CPUs, the ones this process may run on, and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = None
    return {"host_logical_cpus": os.cpu_count(), "usable_cpus": usable, "cpu_model": model}


def copy_probe(crt, args) -> int:
    """Streaming copy of a buffer of known size: measured HBM read+write bandwidth."""
    nbytes = (args.copy_probe // 16) * 16
    src, dst = crt.DeviceArray(np.uint8, nbytes), crt.DeviceArray(np.uint8, nbytes)
    src.fill_bytes(1)
    stream = crt.Stream()
    for _ in range(max(1, args.warmup)):
        crt._lib.check(crt.lib().crt_device_copy(dst.ptr, src.ptr, nbytes, stream.handle))
    stream.sync()
    ev = Events(crt, 2)
    ev.record(0, stream.handle)
    for _ in range(args.steps):
        crt._lib.check(crt.lib().crt_device_copy(dst.ptr, src.ptr, nbytes, stream.handle))
    ev.record(1, stream.handle)
    ms = ev.ms(0, 1) / args.steps
    gbs = 2.0 * nbytes / (ms * 1e-3) / 1e9
    print(json.dumps({"probe": "copy16_kernel", "bytes_read": nbytes, "bytes_written": nbytes,
                      "ms_per_copy": round(ms, 4), "read_plus_write_GBs": round(gbs, 1)}), flush=True)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--format", choices=("stackless", "stack"), default="stackless")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--lds-stack", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline duration")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (profiling runs)")
    ap.add_argument("--spheres", type=int, default=2000)
    ap.add_argument("--skip-batched", action="store_true", help="skip the batched-vs-plain comparison (64 scenes)")
    ap.add_argument("--batched-scenes", type=int, default=64)
    ap.add_argument("--batched-steps", type=int, default=5)
    ap.add_argument("--only", default="", metavar="KERNEL:VARIANT",
                    help="time one kernel choice only (e.g. 1:0 one ray per lane, 3:1 wave walk preset 1, "
                         "w4l2:3:0 the wave walk preset 0 over the BVH4 layout with 2-triangle leaves; a suffix "
                         "+ao:K:V runs the AO launches with kernel K, variant V)")
    ap.add_argument("--no-wide", action="store_true", help="time the BVH2 kernels only")
    ap.add_argument("--no-stack", action="store_true",
                    help="skip the kernels over the stack layout (FlattenStackBVH) of the same meshes")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the large-scene path-tracing workload (15.6M triangles, 3840x2160x4 paths)")
    ap.add_argument("--large-scale", type=int, default=1, metavar="D",
                    help="divide the large scene's terrain grid, sphere count and image size by D (tests)")
    ap.add_argument("--large-preset", type=int, default=-1, metavar="V",
                    help="time only path-kernel preset V on the large scene (default: every preset)")
    ap.add_argument("--large-layout", default="s2", metavar="L",
                    help="the large scene's BVH4: lN = greedy collapse with leaves of up to N triangles, sN = SAH "
                         "collapse (default s2)")
    ap.add_argument("--large-only", action="store_true",
                    help="run only the large-scene path-tracing workload and print its record as one JSON line")
    ap.add_argument("--detail", default="", metavar="PATH",
                    help="also write every candidate's timing and the side workloads' records to PATH (JSON)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="instead of the workload, time the streaming copy kernel over BYTES (HBM probe and "
                         "counter calibration)")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); rays shard across GPUs with no exchange, "
              "so the figure of N GPUs is N independent copies of this run", file=sys.stderr)
        return 2

    import candela_amd as crt
    from candela_amd import scenes

    fmt = crt.FORMAT_STACKLESS if args.format == "stackless" else crt.FORMAT_STACK
    if crt.device_count() < 1:
        raise SystemExit("no HIP device visible")
    if args.copy_probe:
        return copy_probe(crt, args)
    if args.large_only:
        print(json.dumps(time_large(crt, args)), flush=True)
        return 0

    # --- scene + BVH (CPU constructor) -------------------------------------------------
    t0 = time.perf_counter()
    meshes = scenes.icosphere_cloud(count=args.spheres, subdivisions=3, seed=2)
    t1 = time.perf_counter()
    r = crt.RayIntersector(fmt)
    r.add_object(0, meshes)
    t2 = time.perf_counter()
    r.push_entity(0)
    r.buffer_entities()
    r.buffer_data()
    hb = r.host_buffers()
    ntri = int(hb["triangles"].size)
    log(f"scene {ntri} tris built in {t2 - t1:.2f}s (generation {t1 - t0:.2f}s), {hb['nodes'].size} nodes")

    # --- G-buffer and secondary rays --------------------------------------------------
    W, H = 1920, 1080
    iv, ip = scenes.camera_inverses(aspect=W / H, **scenes.CLOUD_CAMERA)
    drays, dhits = r.intersect_primary(W, H, iv, ip)
    crt.lib().crt_device_sync()
    prays, phits = drays.download(), dhits.download()
    pos = hb["vertices"]["position"][:, :3]
    ao, gi = scenes.secondary_rays(prays, phits, pos, hb["triangles"], seed=2)
    n_ao, n_gi = ao.size, gi.size
    log(f"primary hits {n_ao}/{W * H}; {n_ao} AO + {n_gi} GI rays per step")
    d_ao, d_gi = crt.DeviceArray.from_host(ao), crt.DeviceArray.from_host(gi)
    h_ao, h_gi = crt.DeviceArray(crt.HIT_DTYPE, n_ao), crt.DeviceArray(crt.HIT_DTYPE, n_gi)
    scene = r.scene()
    stream = crt.Stream()
    sh = stream.handle

    K = args.steps

    wides = {}

    stack_ctx = {}

    def scene_of(c: dict):
        if c["layout"] == "bvh2":
            return scene
        if c["layout"] == "stack":
            if "r" not in stack_ctx:   # the same meshes through FlattenStackBVH (the reference's stack layout)
                t = time.perf_counter()
                rs = crt.RayIntersector(crt.FORMAT_STACK)
                rs.add_object(0, meshes)
                rs.push_entity(0)
                rs.buffer_entities()
                rs.buffer_data()
                stack_ctx["r"], stack_ctx["scene"] = rs, rs.scene()
                stack_ctx["hb"] = rs.host_buffers()
                log(f"stack layout: {stack_ctx['hb']['nodes'].size} entries, built in {time.perf_counter() - t:.2f}s")
            return stack_ctx["scene"]
        key = wide_key(c)
        if key not in wides:
            t = time.perf_counter()
            wides[key] = crt.upload_wide(fmt, hb, key[0], key[1], scene.triangles, scene.vertices, key[2], 0.0,
                                         SAH_TRI_COST if key[2] else 0.0)
            st = wides[key].host.stats
            log(f"wide layout {key}: {st['node_count']} nodes, fill {st['mean_fill']:.2f}, depth {st['max_depth']}, "
                f"built in {time.perf_counter() - t:.2f}s")
        return wides[key].scene

    def run(c: dict) -> dict:
        """W warm-up + K timed steps of (AO any-hit, GI closest-hit) with one kernel choice."""
        kernel, variant = c["kernel"], c["variant"]
        scene = scene_of(c)
        kw = dict(block_size=c.get("block", args.block), lds_stack=c.get("lds", args.lds_stack), kernel=kernel,
                  variant=variant, stream=stream)
        ak, av = ao_of(c)
        kw_ao = dict(kw, kernel=ak, variant=av, flags=ao_flags(crt, c))
        for _ in range(args.warmup):
            crt.trace(scene, crt.QUERY_ANY, d_ao, h_ao, **kw_ao)
            crt.trace(scene, crt.QUERY_CLOSEST, d_gi, h_gi, **kw)
        stream.sync()
        crt.lib().crt_device_sync()
        ev = Events(crt, 3 * K + 2)
        w0 = time.perf_counter()
        ev.record(0, sh)
        for k in range(K):
            ev.record(1 + 3 * k, sh)
            crt.trace(scene, crt.QUERY_ANY, d_ao, h_ao, **kw_ao)
            ev.record(2 + 3 * k, sh)
            crt.trace(scene, crt.QUERY_CLOSEST, d_gi, h_gi, **kw)
            ev.record(3 + 3 * k, sh)
        ev.record(3 * K + 1, sh)
        stream.sync()
        crt.lib().crt_device_sync()
        w1 = time.perf_counter()
        total_ms = ev.ms(0, 3 * K + 1)
        ao_ms = float(np.mean([ev.ms(1 + 3 * k, 2 + 3 * k) for k in range(K)]))
        gi_ms = float(np.mean([ev.ms(2 + 3 * k, 3 + 3 * k) for k in range(K)]))
        ev.close()
        ms_step = total_ms / K
        res = dict(c, choice=choice_name(c), ms_step=ms_step, wall_ms_step=(w1 - w0) * 1e3 / K, ao_ms=ao_ms,
                   gi_ms=gi_ms, mrays=(n_ao + n_gi) / (ms_step * 1e-3) / 1e6)
        log(f"{choice_name(c)}: {ms_step:.3f} ms/step (wall {res['wall_ms_step']:.3f}), "
            f"AO {ao_ms:.3f} ms, GI {gi_ms:.3f} ms -> {res['mrays']:.1f} Mrays/s")
        return res

    def run_overlapped(c: dict) -> dict:
        """The same step with the AO launch on a second stream: the two launches are independent, so each
        fills the other's tail.  A fork and a join event per step keep every step's launches together; events
        on each stream give each launch's duration inside the timed region.  The records must equal those of a
        sequential step byte for byte."""
        import ctypes as C
        kernel, variant = c["kernel"], c["variant"]
        scene = scene_of(c)
        kw = dict(block_size=c.get("block", args.block), lds_stack=c.get("lds", args.lds_stack), kernel=kernel,
                  variant=variant, stream=stream)
        ak, av = ao_of(c)
        s2 = crt.Stream()
        kw_ao = dict(kw, kernel=ak, variant=av, flags=ao_flags(crt, c), stream=s2)
        lib = crt.lib()
        # a sequential step's records, to compare with
        crt.trace(scene, crt.QUERY_ANY, d_ao, h_ao, **dict(kw_ao, stream=stream))
        crt.trace(scene, crt.QUERY_CLOSEST, d_gi, h_gi, **kw)
        stream.sync()
        seq = (h_ao.download(), h_gi.download())
        h_ao.fill_bytes(0xAB)
        h_gi.fill_bytes(0xAB)
        crt.lib().crt_device_sync()

        def step(ev, k):
            ev.record(2 + 3 * k, sh)                      # fork
            crt._lib.check(lib.crt_stream_wait_event(s2.handle, C.c_void_p(ev.ev[2 + 3 * k])))
            # GI first: its persistent waves take the device, the AO grid's waves fill the slots they leave
            # (0.711 against 0.722 ms per step with AO enqueued first)
            crt.trace(scene, crt.QUERY_CLOSEST, d_gi, h_gi, **kw)
            crt.trace(scene, crt.QUERY_ANY, d_ao, h_ao, **kw_ao)
            ev.record(3 + 3 * k, sh)                      # GI done
            ev.record(4 + 3 * k, s2.handle)               # AO done; join
            crt._lib.check(lib.crt_stream_wait_event(sh, C.c_void_p(ev.ev[4 + 3 * k])))

        warm = Events(crt, 3 * max(1, args.warmup) + 2)
        for k in range(args.warmup):
            step(warm, k)
        stream.sync()
        crt.lib().crt_device_sync()
        warm.close()
        ev = Events(crt, 3 * K + 2)
        w0 = time.perf_counter()
        ev.record(0, sh)
        for k in range(K):
            step(ev, k)
        ev.record(1, sh)
        stream.sync()
        crt.lib().crt_device_sync()
        w1 = time.perf_counter()
        ms_step = ev.ms(0, 1) / K
        gi_ms = float(np.mean([ev.ms(2 + 3 * k, 3 + 3 * k) for k in range(K)]))
        ao_ms = float(np.mean([ev.ms(2 + 3 * k, 4 + 3 * k) for k in range(K)]))
        ev.close()
        s2.close()
        same = h_ao.download().tobytes() == seq[0].tobytes() and h_gi.download().tobytes() == seq[1].tobytes()
        res = dict(c, streams=2, ms_step=ms_step, wall_ms_step=(w1 - w0) * 1e3 / K, ao_ms=ao_ms, gi_ms=gi_ms,
                   ao_ms_alone=c["ao_ms"], gi_ms_alone=c["gi_ms"], records_equal_sequential=same,
                   mrays=(n_ao + n_gi) / (ms_step * 1e-3) / 1e6)
        res["choice"] = choice_name(res)
        log(f"{res['choice']}: {ms_step:.3f} ms/step (wall {res['wall_ms_step']:.3f}) with AO and GI on two streams "
            f"(AO {ao_ms:.3f} ms, GI {gi_ms:.3f} ms from the fork) -> {res['mrays']:.1f} Mrays/s; records equal a "
            f"sequential step: {same}")
        if not same:
            raise RuntimeError("overlapped launches wrote other records than a sequential step")
        return res

    def bvh2(k, v):
        return dict(layout="bvh2", kernel=k, variant=v)

    def stack(k, v, lds=16):
        return dict(layout="stack", kernel=k, variant=v, lds=lds)

    def wide(w, leaf, kernel=crt.KERNEL_WAVE, variant=0, collapse=0):
        return dict(layout=f"wide{w}", width=w, max_leaf=leaf, kernel=kernel, variant=variant, collapse=collapse)

    candidates = [bvh2(crt.KERNEL_THREAD_PER_RAY, 0)]
    if fmt == crt.FORMAT_STACKLESS:
        candidates += [bvh2(crt.KERNEL_WAVE, v) for v in range(6)] + [bvh2(crt.KERNEL_PERSISTENT, v) for v in range(6)]
        candidates += [bvh2(k, v) for k in (crt.KERNEL_SPECULATIVE, crt.KERNEL_SPECULATIVE_REFILL) for v in range(4)]
        if not args.no_wide:
            candidates += [wide(w, 2, crt.KERNEL_THREAD_PER_RAY) for w in (4, 8)]
            candidates += [wide(w, leaf, k, v) for w in (4, 8) for leaf in (2, 4)
                           for k in (crt.KERNEL_WAVE, crt.KERNEL_PERSISTENT) for v in range(10)]
            candidates += [wide(w, 2, k, v) for w in (4, 8)
                           for k in (crt.KERNEL_SPECULATIVE, crt.KERNEL_SPECULATIVE_REFILL) for v in range(6)]
            candidates += [wide(w, leaf, k, v, crt.COLLAPSE_SAH) for w, leaf in ((4, 2), (4, 4), (8, 2))
                           for k in (crt.KERNEL_SPECULATIVE, crt.KERNEL_SPECULATIVE_REFILL) for v in range(6)]
        if not args.no_stack:   # the reference's other layout over the same meshes
            candidates += [stack(crt.KERNEL_THREAD_PER_RAY, 0, lds) for lds in (16, 32)]
            candidates += [stack(k, v, 0) for k in (crt.KERNEL_WAVE, crt.KERNEL_PERSISTENT) for v in range(8)]
            candidates += [stack(k, v, 0) for k in (crt.KERNEL_SPECULATIVE, crt.KERNEL_SPECULATIVE_REFILL)
                           for v in range(4)]
    if args.only:
        main_part, _, ao_part = args.only.replace("+overlap", "").partition("+ao:")
        parts = main_part.split(":")
        if parts[0].startswith("w"):
            sah = "s" in parts[0]
            w, leaf = (int(x) for x in parts[0][1:].split("s" if sah else "l"))
            only = wide(w, leaf, int(parts[1]), int(parts[2]), crt.COLLAPSE_SAH if sah else 0)
        elif parts[0].startswith("s"):
            only = stack(int(parts[1]), int(parts[2]), int(parts[0][1:]))
        else:
            only = bvh2(int(parts[0]), int(parts[1]))
        if ao_part:
            ap = ao_part.split(":")
            only.update(ao_kernel=int(ap[0]), ao_variant=int(ap[1]))
            if ap[2:] == ["s"]:
                only.update(ao_order=1)
        candidates = [only]
    results = [run(c) for c in candidates]
    # The AO and GI launches may use different kernels over the same layout: the fastest GI choice with the
    # fastest AO choice of its layout, timed as one more candidate.
    fastest_gi = min(results, key=lambda r: r["gi_ms"])
    same_layout = [r for r in results if scene_of(r) is scene_of(fastest_gi) and "ao_kernel" not in r]
    fastest_ao = min(same_layout, key=lambda r: r["ao_ms"]) if same_layout else fastest_gi
    if "ao_kernel" not in fastest_gi and \
            (fastest_ao["kernel"], fastest_ao["variant"]) != (fastest_gi["kernel"], fastest_gi["variant"]):
        mixed = {k: v for k, v in fastest_gi.items()
                 if k in ("layout", "width", "max_leaf", "collapse", "kernel", "variant", "lds", "block")}
        mixed.update(ao_kernel=fastest_ao["kernel"], ao_variant=fastest_ao["variant"])
        results.append(run(mixed))
    # child order of the AO walks: the fastest AO choice of the wide layouts again with slot order (the
    # speculative kernels always walk nearest first, so they have no second order to time)
    if not args.only and fastest_gi["layout"] != "bvh2":
        base = min((r for r in results if scene_of(r) is scene_of(fastest_gi)), key=lambda r: r["ms_step"])
        if ao_of(base)[0] not in (crt.KERNEL_SPECULATIVE, crt.KERNEL_SPECULATIVE_REFILL):
            results.append(run(dict({k: v for k, v in base.items() if k in (
                "layout", "width", "max_leaf", "collapse", "kernel", "variant", "lds", "block", "ao_kernel",
                "ao_variant")},
                ao_order=1)))
    thread = results[0]
    # the three fastest choices again with their two launches overlapped on two streams (the best pairing on one
    # stream need not be the best on two)
    top = sorted(results, key=lambda r: -r["mrays"])[:3]
    # ... and the best choice of each of the reference's own layouts, for their per-layout figures
    for lay in ("bvh2", "stack"):
        own = [r for r in results if r["layout"] == lay]
        if own:
            b_ = max(own, key=lambda r: r["mrays"])
            if b_ not in top:
                top.append(b_)
    for c in top:
        results.append(run_overlapped(c))
    best = max(results, key=lambda r: r["mrays"])
    best_scene = scene_of(best)
    best_fmt = best_scene.format
    ms_step = best["ms_step"]
    rays_step = n_ao + n_gi
    value = best["mrays"]

    # --- algorithmic bytes (counting build, untimed) ------------------------------------
    out = {}
    for name, dr, dh, q, n in (("ao", d_ao, h_ao, crt.QUERY_ANY, n_ao), ("gi", d_gi, h_gi, crt.QUERY_CLOSEST, n_gi)):
        cnt = crt.DeviceArray(np.uint32, n)
        k_, v_ = ao_of(best) if q == crt.QUERY_ANY else (best["kernel"], best["variant"])
        crt.trace(best_scene, q, dr, dh, block_size=best.get("block", args.block),
                  lds_stack=best.get("lds", args.lds_stack), kernel=k_, variant=v_,
                  tri_tests=cnt, flags=ao_flags(crt, best) if q == crt.QUERY_ANY else 0)
        crt.lib().crt_device_sync()
        hits = dh.download()
        tests = cnt.download()
        visits = hits["status"] & crt.STATUS_VISITS_MASK
        b = bytes_per_ray(best_fmt, visits, tests)
        out[name] = dict(hits=hits, bytes=float(b.sum()), visits=float(visits.mean()), tests=float(tests.mean()),
                         capped=float(np.mean((hits["status"] & crt.STATUS_ITER_CAP) != 0)),
                         hit_rate=float(np.mean(hits["t"] > 0)))
    gi_s = best["gi_ms"] * 1e-3
    achieved = out["gi"]["bytes"] / gi_s / 1e9
    traffic, traffic_src, counters = measured_traffic(best, args, fmt)
    log(f"GI: {out['gi']['visits']:.1f} node visits, {out['gi']['tests']:.1f} tri tests per ray, "
        f"{out['gi']['bytes'] / n_gi:.0f} B/ray, {achieved:.0f} GB/s algorithmic; "
        f"iteration-capped {out['gi']['capped'] * 100:.2f}%")

    # --- hit attributes (GetData) of the GI hits: one launch, HIP events -----------------------
    attrs = time_attributes(crt, best_scene, h_gi, n_gi, out["gi"]["hits"], stream, K)

    # --- device BVH builder over the same scene; GI rays over its tree (stack layout) ----------
    try:
        gpu_build, gb = time_gpu_build(crt, meshes, stream, K)
        dh_b = crt.DeviceArray(crt.HIT_DTYPE, n_gi)
        crt.trace(gb.scene, crt.QUERY_CLOSEST, d_gi, dh_b, kernel=crt.KERNEL_THREAD_PER_RAY, stream=stream)
        stream.sync()
        ev = Events(crt, 2)
        ev.record(0, sh)
        for _ in range(K):
            crt.trace(gb.scene, crt.QUERY_CLOSEST, d_gi, dh_b, kernel=crt.KERNEL_THREAD_PER_RAY, stream=stream)
        ev.record(1, sh)
        ms_b = ev.ms(0, 1) / K
        ev.close()
        same_t = float(np.mean(dh_b.download()["t"] == out["gi"]["hits"]["t"]))
        sah_ms = thread["gi_ms"] if thread["layout"] == "bvh2" and thread["kernel"] == crt.KERNEL_THREAD_PER_RAY \
            else None
        gpu_build.update(gi_ms_thread_kernel=round(ms_b, 4),
                         gi_ms_thread_kernel_sah=round(sah_ms, 4) if sah_ms is not None else None,
                         gi_t_agreement_with_sah_tree=same_t)
        log(f"gpu-built tree: GI {ms_b:.3f} ms one ray per lane (SAH tree {sah_ms} ms), "
            f"nearest t agreement {same_t:.6f}")
        # the device BVH4 collapse of that tree (wall clock: it reads each level's size back), and the best
        # refilling BVH4 GI kernel over it against the same kernel over the SAH tree's collapse
        gw = crt.GpuWide(gb, 4, stream=stream)
        w0 = time.perf_counter()
        for _ in range(K):
            gw.build(stream)
        collapse_ms = (time.perf_counter() - w0) * 1e3 / K
        k_w, v_w = (best["kernel"], best["variant"]) if best["layout"] == "wide4" else (crt.KERNEL_PERSISTENT, 8)
        dh_w = crt.DeviceArray(crt.HIT_DTYPE, n_gi)
        crt.trace(gw.scene, crt.QUERY_CLOSEST, d_gi, dh_w, kernel=k_w, variant=v_w, stream=stream)
        stream.sync()
        ev = Events(crt, 2)
        ev.record(0, sh)
        for _ in range(K):
            crt.trace(gw.scene, crt.QUERY_CLOSEST, d_gi, dh_w, kernel=k_w, variant=v_w, stream=stream)
        ev.record(1, sh)
        ms_w = ev.ms(0, 1) / K
        ev.close()
        same_w = float(np.mean(dh_w.download()["t"] == out["gi"]["hits"]["t"]))
        gpu_build.update(wide_collapse_ms_wall=round(collapse_ms, 3), wide_nodes=int(gw.node_count),
                         gi_ms_wide_kernel=round(ms_w, 4), gi_wide_kernel=f"{k_w}:{v_w}",
                         gi_t_agreement_wide=same_w)
        log(f"device BVH4 collapse: {gw.node_count} nodes in {collapse_ms:.3f} ms (wall); GI {ms_w:.3f} ms with "
            f"kernel {k_w}:{v_w} over it, nearest t agreement {same_w:.6f}")
        # the asynchronous collapse (nothing read back): 64 levels, then sized to the levels the tree needs
        # plus 2 (what a per-frame rebuild would launch); the whole per-frame rebuild (build, collapse, AO and
        # GI traces) captured once into a graph and replayed K times
        def time_async(levels):
            g_ = crt.GpuWide(gb, 4, stream=stream, max_levels=levels)
            stream.sync()
            ev_ = Events(crt, 2)
            ev_.record(0, sh)
            for _ in range(K):
                g_.build(stream)
            ev_.record(1, sh)
            ms_ = ev_.ms(0, 1) / K
            ev_.close()
            return g_, ms_, g_.result()

        ga64, ms_async64, _ = time_async(64)
        need = ga64.levels
        del ga64
        ga, ms_async, (n_async, status) = time_async(need + 2)
        dh_ao = crt.DeviceArray(crt.HIT_DTYPE, n_ao)
        for _ in range(2):
            crt.trace(ga.scene, crt.QUERY_ANY, d_ao, dh_ao, kernel=k_w, variant=v_w, stream=stream)
            crt.trace(ga.scene, crt.QUERY_CLOSEST, d_gi, dh_w, kernel=k_w, variant=v_w, stream=stream)
        stream.sync()
        graph = crt.Graph(stream)
        q_ao, q_gi = crt.queue_block(), crt.queue_block()   # the graph owns its work-queue counters
        with graph:
            gb.build(stream)
            ga.build(stream)
            crt.trace(ga.scene, crt.QUERY_ANY, d_ao, dh_ao, kernel=k_w, variant=v_w, stream=stream, queue=q_ao)
            crt.trace(ga.scene, crt.QUERY_CLOSEST, d_gi, dh_w, kernel=k_w, variant=v_w, stream=stream, queue=q_gi)
        graph.launch()
        stream.sync()
        ev = Events(crt, 2)
        ev.record(0, sh)
        for _ in range(K):
            graph.launch()
        ev.record(1, sh)
        ms_frame = ev.ms(0, 1) / K
        ev.close()
        graph.close()
        same_frame = float(np.mean(dh_w.download()["t"] == out["gi"]["hits"]["t"]))
        gpu_build.update(wide_collapse_async_ms_64_levels=round(ms_async64, 3), wide_levels=need,
                         wide_collapse_async_ms=round(ms_async, 3), wide_collapse_async_levels=need + 2,
                         wide_collapse_async_ok=bool(status == 0 and n_async == gw.node_count),
                         rebuild_frame_graph_ms=round(ms_frame, 3),
                         rebuild_frame_mrays_s=round((n_ao + n_gi) / (ms_frame * 1e-3) / 1e6, 1),
                         rebuild_frame_gi_t_agreement=same_frame)
        log(f"async BVH4 collapse: {ms_async64:.3f} ms at 64 levels, {ms_async:.3f} ms at {need + 2} (the tree "
            f"needs {need}; status {status}, {n_async} nodes); rebuild frame in one "
            f"graph (build + collapse + AO + GI): {ms_frame:.3f} ms, nearest t agreement {same_frame:.6f}")
        del gb, dh_b, gw, dh_w, ga, dh_ao
    except Exception as e:   # reported, never fatal to the headline line
        gpu_build = {"error": f"{type(e).__name__}: {e}"}
        log(f"gpu build failed: {gpu_build['error']}")

    # --- wavefront ray generation from the primary G-buffer: one launch chain, HIP events -------
    raygen = time_raygen(crt, scene, drays, dhits, prays, phits, hb, stream, K)

    # --- config 4: differentiable shading, forward and backward passes ------------------------
    shading = None
    if not args.only:
        try:
            shading = time_shading(crt, stream, K)
        except Exception as e:   # reported, never fatal to the headline line
            shading = {"error": f"{type(e).__name__}: {e}"}
            log(f"shading failed: {shading['error']}")

    # --- large scene: multi-bounce paths with in-kernel regeneration --------------------------
    large = None
    if not args.no_large and not args.only:
        try:
            large = time_large(crt, args)
        except Exception as e:   # reported, never fatal to the headline line
            large = {"error": f"{type(e).__name__}: {e}"}
            log(f"large scene failed: {large['error']}")

    # --- batched persistent kernel: config-2 batches and 64 heterogeneous scenes ---------------
    batched = None
    c2_batched = None
    if not args.skip_batched and fmt == crt.FORMAT_STACKLESS:
        def best_of(pred):
            rs = [r for r in results if pred(r)]
            return max(rs, key=lambda r: r["mrays"]) if rs else None

        plain_choice = {}
        b2 = best_of(lambda r: r["layout"] == "bvh2")
        if b2 is None:   # --only named a wide layout: the BVH2 plain launches take the wave kernel, preset 1
            b2 = bvh2(crt.KERNEL_WAVE, 1)
        plain_choice[None] = (b2["kernel"], b2["variant"], *ao_of(b2), 0)
        for w in (4, 8):
            bw = best_of(lambda r: r["layout"] == f"wide{w}" and r["max_leaf"] == 2 and not r.get("collapse"))
            if bw is not None:
                plain_choice[(w, 2)] = (bw["kernel"], bw["variant"], *ao_of(bw), bw.get("ao_order", 0))
        bs = best_of(lambda r: r["layout"] == "wide4" and r["max_leaf"] == 2 and r.get("collapse"))
        if bs is not None:   # the SAH-collapsed BVH4 with 2-triangle leaves
            plain_choice[(4, 2, crt.COLLAPSE_SAH)] = (bs["kernel"], bs["variant"], *ao_of(bs), bs.get("ao_order", 0))
        c2_batched = time_batched(crt, [(best_scene, crt.QUERY_ANY, d_ao, h_ao),
                                        (best_scene, crt.QUERY_CLOSEST, d_gi, h_gi)],
                                  K, args.warmup, stream, wide=best["layout"] != "bvh2")
        log(f"config-2 batches in one persistent launch: {c2_batched['ms']:.3f} ms/step -> "
            f"{rays_step / (c2_batched['ms'] * 1e-3) / 1e6:.1f} Mrays/s ({c2_batched['config']})")
        batched = batched_vs_plain(crt, args, plain_choice, stream)

    # --- CPU baseline on a bounded sample --------------------------------------------------
    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle
        try:
            usable = len(os.sched_getaffinity(0))
        except (AttributeError, OSError):
            usable = os.cpu_count() or 1
        threads = max(1, min(args.cpu_threads, usable))
        nodes, tris, verts, ents = hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"]
        probe = 4096
        ts = time.perf_counter()
        oracle.trace(fmt, nodes, tris, verts, ents, 0, gi[:probe], threads=threads)
        per_ray = (time.perf_counter() - ts) / probe
        # all cores: one warm-up pass, then the best of 3 timed passes over a sample sized to about
        # cpu_seconds in total; one thread: the same over a sample a `threads`-th of that size
        n_s = int(min(n_gi, max(probe, args.cpu_seconds / 5 / 2 / max(per_ray, 1e-9))))

        def spread(a, ng):
            """ng rays at an even stride over the whole image (rays are in pixel order)."""
            return np.ascontiguousarray(a[np.linspace(0, a.size - 1, ng).astype(np.int64)])

        def best_of_3(ng, nthreads):
            rates, hits = [], None
            sg, sa = spread(gi, ng), spread(ao, ng)
            for rep in range(4):
                ts = time.perf_counter()
                hg = oracle.trace(fmt, nodes, tris, verts, ents, 0, sg, threads=nthreads)
                ha = oracle.trace(fmt, nodes, tris, verts, ents, 1, sa, threads=nthreads)
                dt = time.perf_counter() - ts
                if rep > 0:
                    rates.append(2 * ng / dt / 1e6)
                hits = (hg, ha)
            return max(rates), hits

        cpu_value, _ = best_of_3(n_s, threads)
        n_1 = max(probe // 4, n_s // threads)
        cpu_1, _ = best_of_3(n_1, 1)
        host = host_cpu_info()
        cpu = {"value": round(cpu_value, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "value_1thread": round(cpu_1, 4), **host,
               "sample": f"{n_s} GI closest-hit + {n_s} RTAO any-hit rays taken at an even stride over the whole "
                         f"image, on {threads} worker threads; {n_1} + {n_1} on 1 thread; best of 3 passes after "
                         f"1 warm-up (C restatement of the GLSL walk, -O3); BVH build excluded "
                         f"({round(t2 - t1, 2)} s)"}
        # parity over every ray of the step
        g, a = out["gi"]["hits"], out["ao"]["hits"]
        ohg = oracle.trace(fmt, nodes, tris, verts, ents, 0, gi, threads=threads)
        oha = oracle.trace(fmt, nodes, tris, verts, ents, 1, ao, threads=threads)
        def bvh2_twin(stack_layout, hb_, q, rays_, kernel, variant, walk):
            # the speculative kernels' twin replays the preset's schedule (a ray's record does not depend on
            # which rays share its wave, so this holds for the refilling launch too); the other kernels' records
            # are the reference walk's (`walk`, computed on demand)
            if kernel in (crt.KERNEL_SPECULATIVE, crt.KERNEL_SPECULATIVE_REFILL):
                fn = oracle.trace_stack_spec if stack_layout else oracle.trace_stackless_spec
                presets = oracle.SBVH2_PRESETS if kernel == crt.KERNEL_SPECULATIVE else oracle.SBVH2_P_PRESETS
                return fn(hb_["nodes"], hb_["triangles"], hb_["vertices"], hb_["entities"], q, rays_,
                          schedule=presets[variant], threads=threads)
            return walk()

        if best["layout"] in ("bvh2", "stack"):
            on_stack = best["layout"] == "stack" or fmt == crt.FORMAT_STACK
            hb_ = stack_ctx["hb"] if best["layout"] == "stack" else dict(nodes=nodes, triangles=tris,
                                                                          vertices=verts, entities=ents)

            def walk(q, rays_, base):
                if best["layout"] == "bvh2":
                    return base
                return oracle.trace(crt.FORMAT_STACK, hb_["nodes"], hb_["triangles"], hb_["vertices"],
                                    hb_["entities"], q, rays_, threads=threads)

            twin_g = bvh2_twin(on_stack, hb_, 0, gi, best["kernel"], best["variant"], lambda: walk(0, gi, ohg))
            twin_a = bvh2_twin(on_stack, hb_, 1, ao, *ao_of(best), lambda: walk(1, ao, oha))
        else:   # the host twin of the wide walk, on the same wide layout
            w = wides[wide_key(best)].host

            def wide_twin(q, rays_, kernel, variant, flags):
                # the speculative kernels' twin replays the preset's schedule, for the refilling launch too
                if kernel in (crt.KERNEL_SPECULATIVE, crt.KERNEL_SPECULATIVE_REFILL):
                    presets = oracle.SPEC_PRESETS if kernel == crt.KERNEL_SPECULATIVE else oracle.SPEC_P_PRESETS
                    return oracle.trace_wide_spec(w.width, w.groups, w.records, w.entities, q, rays_,
                                                  schedule=presets[variant], flags=flags, threads=threads)
                return oracle.trace_wide(w.width, w.groups, w.records, w.entities, q, rays_, threads=threads,
                                         flags=flags)

            twin_g = wide_twin(0, gi, best["kernel"], best["variant"], 0)
            twin_a = wide_twin(1, ao, *ao_of(best), ao_flags(crt, best))
        gi_miss = g["tri"] != ohg["tri"]
        parity = {"rays": int(g.size + a.size),
                  # bit-identical records against the twin of each launch's kernel
                  "exact": float((np.concatenate([g, a]).tobytes() == np.concatenate([twin_g, twin_a]).tobytes())),
                  "gi_exact_vs_twin": bool(g.tobytes() == twin_g.tobytes()),
                  "ao_exact_vs_twin": bool(a.tobytes() == twin_a.tobytes()),
                  "gi_t_equal_to_twin": float(np.mean(g["t"] == twin_g["t"])),
                  "ao_occlusion_equal_to_twin": float(np.mean((a["t"] > 0) == (twin_a["t"] > 0))),
                  # closest hits: the triangle is unique up to exact distance ties (shared edges and vertices)
                  "gi_id_agreement_vs_reference_walk": float(1.0 - gi_miss.mean()),
                  "gi_id_mismatches": int(gi_miss.sum()),
                  "gi_id_mismatches_at_equal_t": int((gi_miss & (g["t"] == ohg["t"])).sum()),
                  "gi_t_agreement_vs_reference_walk": float(np.mean(g["t"] == ohg["t"])),
                  # any hits: occluded or not is fixed, which occluder is found first depends on the walk order
                  "ao_occlusion_agreement_vs_reference_walk": float(np.mean((a["t"] > 0) == (oha["t"] > 0))),
                  "ao_first_occluder_same_as_reference_walk": float(np.mean(a["tri"] == oha["tri"]))}
        log(f"cpu baseline {cpu_value:.3f} Mrays/s on {threads} threads, {cpu_1:.3f} Mrays/s on 1 thread; "
            f"parity {parity}")

    def short_choice(c):
        return None if c is None else {"choice": choice_name(c), "mrays_s": round(c["mrays"], 2),
                                       "ms_step": round(c["ms_step"], 4), "gi_ms": round(c["gi_ms"], 4),
                                       "ao_ms": round(c["ao_ms"], 4)}

    def best_where(pred):
        rs = [r for r in results if pred(r)]
        return max(rs, key=lambda r: r["mrays"]) if rs else None

    per_format = {
        "stackless": short_choice(best_where(lambda r: r["layout"] == "bvh2" and fmt == crt.FORMAT_STACKLESS)),
        "stack": short_choice(best_where(lambda r: (r["layout"] == "bvh2" and fmt == crt.FORMAT_STACK)
                                         or r["layout"] == "stack")),
        "bvh4": short_choice(best_where(lambda r: r["layout"] == "wide4")),
        "bvh8": short_choice(best_where(lambda r: r["layout"] == "wide8")),
    }
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": gi_choice_name(best) + " (GI closest hit)", "launch_ms": round(gi_s * 1e3, 4),
                "bytes_per_launch": out["gi"]["bytes"], "bytes_per_ray": round(out["gi"]["bytes"] / n_gi, 1),
                "traffic_source": traffic_src}
    if best.get("streams") == 2:
        # launch_ms is the GI launch's duration inside the overlapped timed region (sharing the GPU with the
        # AO launch); the same launch alone, as timed in the sequential step:
        roofline["launch_ms_alone"] = round(best["gi_ms_alone"], 4)
        roofline["achieved_alone"] = round(out["gi"]["bytes"] / (best["gi_ms_alone"] * 1e-3) / 1e9, 1)
    if traffic:
        # algorithmic bytes price every node and record fetch as HBM bytes; the caches serve most of them
        roofline["traffic_frac_of_peak"] = round(traffic / gi_s / 1e9 / HBM_PEAK_GBS, 4)
        roofline.update(counters)
    full = {
        "kernels": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()} for r in results],
        "kernel_label": kernel_label(best),
        "batched_vs_plain": batched,
        "c2_batched": c2_batched,
        "attributes": attrs,
        "raygen": raygen,
        "shading": shading,
        "gpu_build": gpu_build,
        "large_scene": large,
        "cpu_baseline": cpu,
        "parity": parity,
        "detail": {k: {kk: vv for kk, vv in v.items() if kk != "hits"} for k, v in out.items()},
    }
    log("detail " + json.dumps(full))
    if args.detail:
        with open(args.detail, "w") as f:
            json.dump(full, f, indent=1)

    def pick(d, keys):
        return None if not isinstance(d, dict) else {k: d[k] for k in keys if k in d}

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "incoherent GI closest-hit + RTAO any-hit rays from a 1920x1080 G-buffer, "
                        "2000 icospheres (2.56M tris)",
            "triangles": ntri,
            "rays_per_step": rays_step,
            "kernel": choice_name(best),
            "parallelism": "single GPU",
            "streams": best.get("streams", 1),
        },
        "plain_by_format": per_format,
        "batched_mrays_s": (round(rays_step / (c2_batched["ms"] * 1e-3) / 1e6, 2) if c2_batched else None),
        "batched_vs_plain": pick(batched, ("plain_launches", "plain_ms", "plain_layout", "batched_ms",
                                           "batched_layout", "batched_mrays_s", "speedup", "speedup_same_layout",
                                           "hits_identical", "timed_plain_t_and_occlusion_equal",
                                           "library_autotuner")),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
        "large_scene": pick(large, ("triangles", "paths", "segments", "ms_per_frame", "mrays_s", "parity",
                                    "error")),
        "gpu_build": pick(gpu_build, ("triangles", "ms", "wide_collapse_async_ms", "rebuild_frame_graph_ms",
                                      "error")),
        "attributes_mhits_s": attrs.get("mhits_s") if attrs else None,
        "raygen_mrays_s": raygen.get("mrays_s") if raygen else None,
        "shading": pick(shading, ("forward_ms", "backward_ms", "forward_mrays_s", "backward_mrays_s", "error")),
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
