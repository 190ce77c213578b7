// One wave traces up to 64 rays (one per lane) through a stackless BVH with
// leaf batching and a K-node fetch window.  Used by the grid, persistent and
// batched launchers (persistent.hip, batched.hip).
//
// Per ray the operations are those of IntersectBVHStackless /
// IntersectBVHStacklessOcclusion inside the IntersectScene entity loop
// (Source/Core/Shaders/Intersectors/Include/TraverseBVHStackless.glsl:175-366,
// :463-579), in the same order, so results are bit-identical to the host
// oracle.  The wave-level schedule differs:
//
//   node phase  every lane that is not parked steps through nodes; one memory
//               round trip fetches a window of K consecutive nodes (K*32
//               bytes) and the lane keeps stepping while its next node lies in
//               the window -- in the pre-order layout the next node is the
//               successor after an inner-node hit and after a leaf whose right
//               sibling follows it, so short runs cost one fetch;
//   leaf phase  a lane whose box test enters a leaf parks; parked leaves are
//               tested together once LEAF_BATCH lanes wait, after INNER node
//               steps or when no lane can step (the leaf must be tested before
//               the lane's next box test, because a hit shortens TMax);
//   walk end    next entity, or the ray's hit record.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"
#include "device_math.h"
#include "trace_launch.h"

namespace crt {
namespace pdev {

using namespace ::crt::dm;

// Per-lane state of one ray in the scene loop.
//
// The walk writes its accepted hits straight into the scene-level record: the
// walk starts from the scene loop's TMax and only accepts t below it, so every
// hit the walk keeps also wins at scene level (the reference copies the walk's
// result when `T > 0 && T < TMax`, TraverseBVHStackless.glsl:294-300, which
// then always holds).  tmax is the walk's TMax = the scene loop's TMax for a
// closest-hit query, the ray's tmax for an any-hit query.
struct Lane {
    int32_t slot;      // ray index (a launch holds < 2^31 rays)
    f3 lro, lrd, inv;  // ray in the current entity's space
    float ray_tmax;
    int32_t ent;       // current entity
    int32_t start, count;
    int32_t ptr, it;
    float tmax;        // TMax of the walk (closest: shrinks with every hit)
    int32_t tri, mesh, hit_ent;
    uint32_t visits, status, tests;
    int32_t leaf_packed, leaf_link;
};

// The query is ANY, or with DYN the member dyn_any, set before run (one
// compiled walk serving both queries, for kernels that mix them: two inlined
// walks would add up their registers).
template <bool ANY, int INNER, int LEAF_BATCH, int K, bool DYN = false>
struct WaveWalk {
    static_assert(K >= 1 && K <= 4, "window of 1 to 4 nodes");

    Lane L;
    bool active, parked, walk_done;
    bool dyn_any;
    __device__ __forceinline__ bool any() const { return DYN ? dyn_any : ANY; }

    // Next entity that takes part (IgnoreTransparent skips alpha < 0.99);
    // returns false when the entity loop is over.
    __device__ __forceinline__ bool begin_walk(const SceneArgs& s, const crt_ray* __restrict__ rays, uint32_t flags) {
        int32_t e = L.ent + 1;
        while (e < s.entity_count && (flags & CRT_FLAG_IGNORE_TRANSPARENT) &&
               __int_as_float(s.entities[e].data[1]) < 0.99f)
            ++e;
        L.ent = e;
        if (e >= s.entity_count) return false;
        const crt_entity& en = s.entities[e];
        const float4* rp = reinterpret_cast<const float4*>(rays + L.slot);
        const float4 a = rp[0], b = rp[1];
        L.lro = xform(en.inverse, mk(a.x, a.y, a.z), 1.0f);
        L.lrd = xform(en.inverse, mk(b.x, b.y, b.z), 0.0f);
        L.inv = mk(1.0f / L.lrd.x, 1.0f / L.lrd.y, 1.0f / L.lrd.z);
        L.start = en.node_offset;
        L.count = en.node_count;
        L.ptr = L.start;
        L.it = 0;
        if (any()) L.tmax = L.ray_tmax;
        walk_done = L.ptr < 0;  // `while (Pointer >= 0 ...)` never runs
        return true;
    }

    __device__ __forceinline__ void finish(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                           crt_hit* __restrict__ hits, uint32_t* __restrict__ tri_tests,
                                           uint32_t flags) {
        float4 tuvw = make_float4(-1.0f, -1.0f, -1.0f, -1.0f);
        const int32_t tri = L.tri, mesh = L.mesh, ent = L.hit_ent;
        // ClosestT > 0 exactly when a triangle was accepted; its t is then L.tmax
        const float closest = tri >= 0 ? L.tmax : -1.0f;
        if (any()) {
            if (tri >= 0) tuvw.x = closest;
        } else {
            const bool ok = closest > 0.0f && ((flags & CRT_FLAG_TRI0_IS_MISS) ? tri > 0 : tri >= 0);
            if (ok) {
                const crt_entity& e = s.entities[ent];
                const float4* rp = reinterpret_cast<const float4*>(rays + L.slot);
                const float4 a = rp[0], b = rp[1];
                const f3 lro = xform(e.inverse, mk(a.x, a.y, a.z), 1.0f);
                const f3 lrd = xform(e.inverse, mk(b.x, b.y, b.z), 0.0f);
                const int4 tr = *reinterpret_cast<const int4*>(s.tris[tri].packed);
                const f3 va = vertex_pos(s.verts, tr.x), vb = vertex_pos(s.verts, tr.y), vc = vertex_pos(s.verts, tr.z);
                const f3 p = mk(lro.x + lrd.x * closest, lro.y + lrd.y * closest, lro.z + lrd.z * closest);
                const f3 uvw = barycentrics(p, va, vb, vc);
                tuvw = make_float4(closest, uvw.x, uvw.y, uvw.z);
            }
        }
        float4* o = reinterpret_cast<float4*>(hits + L.slot);
        o[0] = tuvw;
        o[1] = make_float4(__int_as_float(tri), __int_as_float(mesh), __int_as_float(ent),
                           __uint_as_float(pack_status(L.visits, L.status)));
        if (tri_tests) tri_tests[L.slot] = L.tests;
    }

    __device__ __forceinline__ bool test_one(int32_t idx, int4 tr, f3 a, f3 b, f3 c) {
        const float t = tri_t(L.lro, L.lrd, a, b, c);
        L.tests++;
        if (t > 0.0f && t < L.tmax) {
            L.tmax = t;
            L.mesh = tr.w;
            L.tri = idx;
            L.hit_ent = L.ent;
            if (any()) return true;
        }
        return false;
    }

    // One node of the walk from (nmin, nmax) already fetched for L.ptr.
    // Returns true when the walk continues to node L.ptr + 1 (an inner node
    // was entered), the only successor a fetch window can already hold.
    __device__ __forceinline__ bool node_step(float4 nmin, float4 nmax) {
        L.it++;
        const float tb = bounds_stackless(nmin, nmax, L.lro, L.inv, L.tmax);
        const int32_t link = __float_as_int(nmax.w);
        const int32_t packed = __float_as_int(nmin.w);
        if (tb > 0.0f && tb < L.tmax) {
            if (packed != -1) {
                parked = true;
                L.leaf_packed = packed;
                L.leaf_link = link;
                return false;
            }
            L.ptr++;
            return true;
        }
        if (link < 0) {
            walk_done = true;
            return false;
        }
        const int32_t next = link + L.start;
        const bool succ = next == L.ptr + 1;   // a leaf's right sibling follows it
        L.ptr = next;
        return succ;
    }

    // Loop-top checks of the walk: iteration cap, then the index guard.
    __device__ __forceinline__ bool may_step(const SceneArgs& s, int32_t max_iter) {
        if (L.it >= max_iter) {
            L.status |= CRT_STATUS_ITER_CAP;
            walk_done = true;
            return false;
        }
        if (L.ptr < L.start || L.ptr > L.start + L.count || int64_t(L.ptr) >= s.total_nodes) {
            L.status |= CRT_STATUS_BAD_INDEX;
            walk_done = true;
            return false;
        }
        return true;
    }

    __device__ void run(const SceneArgs& s, const crt_ray* __restrict__ rays, crt_hit* __restrict__ hits,
                        int64_t slot, uint32_t flags, int32_t max_iter, uint32_t* __restrict__ tri_tests) {
        const float4* __restrict__ nodes = reinterpret_cast<const float4*>(s.nodes);
        active = slot >= 0;
        parked = false;
        walk_done = false;
        if (active) {
            L.slot = int32_t(slot);
            const float4 a = reinterpret_cast<const float4*>(rays + slot)[0];
            L.ray_tmax = a.w;
            L.ent = -1;
            L.tmax = a.w;
            L.tri = L.mesh = L.hit_ent = -1;
            L.visits = L.status = L.tests = 0;
            if (!begin_walk(s, rays, flags)) {
                finish(s, rays, hits, tri_tests, flags);
                active = false;
            }
        }
        while (__ballot(active)) {
            // --- node phase ------------------------------------------------------
            for (int k = 0; k < INNER; ++k) {
                if (active && !parked && !walk_done && may_step(s, max_iter)) {
                    // one round trip fetches K consecutive nodes; the lane steps
                    // on while its next node is the following one
                    float4 wn[2 * K];
                    const int64_t p0 = L.ptr;
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        if (j == 0 || p0 + j < s.total_nodes) {
                            wn[2 * j] = nodes[2 * (p0 + j)];
                            wn[2 * j + 1] = nodes[2 * (p0 + j) + 1];
                        }
                    }
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        if (j > 0 && !may_step(s, max_iter)) break;
                        if (!node_step(wn[2 * j], wn[2 * j + 1])) break;
                    }
                }
                const unsigned long long stepping = __ballot(active && !parked && !walk_done);
                if (stepping == 0 || __popcll(__ballot(parked)) >= LEAF_BATCH) break;
            }

            // --- leaf phase ------------------------------------------------------
            if (__ballot(parked)) {
                if (parked) {
                    const int32_t first = L.leaf_packed >> 4;
                    const int32_t len = L.leaf_packed & 0xF;
                    bool done = false;
                    if (len > 0) {
                        // both triangles of a two-triangle leaf are fetched before either test
                        const int4 t0 = *reinterpret_cast<const int4*>(s.tris[first].packed);
                        int4 t1 = t0;
                        if (len > 1) t1 = *reinterpret_cast<const int4*>(s.tris[first + 1].packed);
                        const f3 a0 = vertex_pos(s.verts, t0.x), b0 = vertex_pos(s.verts, t0.y),
                                 c0 = vertex_pos(s.verts, t0.z);
                        f3 a1 = a0, b1 = b0, c1 = c0;
                        if (len > 1) {
                            a1 = vertex_pos(s.verts, t1.x);
                            b1 = vertex_pos(s.verts, t1.y);
                            c1 = vertex_pos(s.verts, t1.z);
                        }
                        done = test_one(first, t0, a0, b0, c0);
                        if (!done && len > 1) done = test_one(first + 1, t1, a1, b1, c1);
                        for (int32_t idx = first + 2; !done && idx < first + len; ++idx) {
                            const int4 tr = *reinterpret_cast<const int4*>(s.tris[idx].packed);
                            done = test_one(idx, tr, vertex_pos(s.verts, tr.x), vertex_pos(s.verts, tr.y),
                                            vertex_pos(s.verts, tr.z));
                        }
                    }
                    parked = false;
                    if (done || L.leaf_link < 0)
                        walk_done = true;
                    else
                        L.ptr = L.leaf_link + L.start;
                }
            }

            // --- walk ends: next entity or the ray's result ----------------------
            // the ends run once kWalkEndBatch lanes wait or no lane can step
            const unsigned long long ends = __ballot(walk_done);
            if (ends && (__popcll(ends) >= kWalkEndBatch || !__ballot(active && !parked && !walk_done))) {
                if (walk_done) {
                    walk_done = false;
                    L.visits += uint32_t(L.it);
                    // any-hit: the first entity with a hit ends the scene loop
                    bool fin = any() && L.tri >= 0;
                    if (!fin) fin = !begin_walk(s, rays, flags);
                    if (fin) {
                        finish(s, rays, hits, tri_tests, flags);
                        active = false;
                    }
                }
            }
        }
    }
};

}  // namespace pdev
}  // namespace crt
