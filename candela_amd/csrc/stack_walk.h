// Wave-scheduled walk of the reference's stack layout (FlattenStackBVH,
// BVHConstructor.cpp:847-930): one ray per lane, leaves parked and tested by
// the wave together, used by the grid and refilling launchers of stack.hip.
//
// Per ray and entity the operations are those of IntersectBVHStack /
// IntersectBVHStackOcclusion (Include/TraverseBVHStack.glsl:168-324,
// :509-657) inside the IntersectScene entity loop, in the same order, so the
// hit records, node visits and triangle tests equal the one-ray-per-lane
// walk (thread_walk.h: walk_stack) and the host oracle bit for bit:
//
//   guards    iteration cap (CRT_STATUS_ITER_CAP), then the entry index
//             within [start, start + count] (CRT_STATUS_BAD_INDEX)
//   node      one 64-byte entry holds both children.  Inner children get the
//             stack include's RayBounds (:107-116) with the tmax the node was
//             entered with; leaf children are not box-tested, their triangles
//             are tested left leaf first, then right.  The next entry is
//             decided from the two box distances alone: both hit, the nearer
//             (the right one only if strictly nearer) and the other pushed --
//             a push at 63 entries ends the walk with
//             CRT_STATUS_STACK_OVERFLOW, after that node's leaves; one hit,
//             that one; none, a pop, or the end of the walk
//   schedule  a lane whose node has leaves parks; once LEAF_BATCH lanes are
//             parked (or no lane can step) every parked lane tests its
//             triangles, two per fetch, before its next node step -- whose box
//             tests need the tmax those tests leave
//
// The stack keeps its first LDS_DEPTH entries in LDS (entry-major, lane i of
// a wave on bank i) and the rest in a per-lane scratch array.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"
#include "device_math.h"
#include "queue.h"
#include "trace_launch.h"

namespace crt {
namespace sdev {

using namespace ::crt::dm;

// A push at this many entries aborts the walk (TraverseBVHStack.glsl:293-295).
constexpr int32_t kStackCap = 63;

template <int LDS_DEPTH>
constexpr int spill_entries() {
    return kStackCap - LDS_DEPTH > 0 ? kStackCap - LDS_DEPTH : 1;
}

// One lane's stack: LDS column (stride ints apart) then per-lane scratch.
template <int LDS_DEPTH>
struct NodeStack {
    int32_t* lds;
    int32_t stride;
    int32_t* spill;

    __device__ __forceinline__ LdsInt* col() const { return (LdsInt*)lds; }   // see LdsInt
    __device__ __forceinline__ void put(int32_t sp, int32_t v) {
        if (sp < LDS_DEPTH)
            col()[sp * stride] = v;
        else
            spill[sp - LDS_DEPTH] = v;
    }
    __device__ __forceinline__ int32_t get(int32_t sp) const {
        return sp < LDS_DEPTH ? col()[sp * stride] : spill[sp - LDS_DEPTH];
    }
};

enum StepResult : int32_t { kNext = 0, kLeaves = 1, kEnd = 2 };

__device__ __forceinline__ bool leaf_nonempty(int32_t packed) { return packed != -1 && (packed & 0xF) > 0; }

template <bool ANY, int LDS_DEPTH, int INNER, int LEAF_BATCH>
struct StackWave {
    f3 ro, rd, inv;        // ray in the current entity's space
    float tmax;            // walk TMax: closest shrinks with every hit; any-hit starts each walk at ray_tmax
    float ray_tmax;
    int32_t slot, ent, hit_ent;
    int32_t start, count, cur, sp, it;
    int32_t tri, mesh;
    uint32_t tests, status, visits;
    int32_t leaf0, leaf1;  // the last node's leaf children (packed start << 4 | count, or -1)
    int32_t leaf, off;     // leaf under test, its triangles done
    int32_t pi;            // 1: leaf1 still pending, 2: none
    uint32_t end_bits;     // the walk ends after its pending leaves (kEndFlag), with these status bits
    bool active, leafing, walk_done;

    static constexpr uint32_t kEndFlag = 1u << 31;

    // Next entity that takes part (IgnoreTransparent skips alpha < 0.99);
    // false when the entity loop is over.
    __device__ __forceinline__ bool begin_walk(const SceneArgs& s, const crt_ray* __restrict__ rays, uint32_t flags) {
        int32_t e = ent + 1;
        while (e < s.entity_count && (flags & CRT_FLAG_IGNORE_TRANSPARENT) &&
               __int_as_float(s.entities[e].data[1]) < 0.99f)
            ++e;
        ent = e;
        if (e >= s.entity_count) return false;
        const crt_entity& en = s.entities[e];
        const float4* rp = reinterpret_cast<const float4*>(rays + slot);
        const float4 a = rp[0], b = rp[1];
        ro = xform(en.inverse, mk(a.x, a.y, a.z), 1.0f);
        rd = xform(en.inverse, mk(b.x, b.y, b.z), 0.0f);
        inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
        start = en.node_offset;
        count = en.node_count;
        cur = start;
        sp = 0;
        it = 0;
        end_bits = 0;
        if (ANY) tmax = ray_tmax;
        walk_done = false;
        return true;
    }

    __device__ __forceinline__ void finish(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                           crt_hit* __restrict__ hits, uint32_t* __restrict__ tri_tests,
                                           uint32_t flags) {
        float4 tuvw = make_float4(-1.0f, -1.0f, -1.0f, -1.0f);
        // ClosestT > 0 exactly when a triangle was accepted; its t is then tmax
        const float closest = tri >= 0 ? tmax : -1.0f;
        if (ANY) {
            if (tri >= 0) tuvw.x = closest;
        } else {
            const bool ok = closest > 0.0f && ((flags & CRT_FLAG_TRI0_IS_MISS) ? tri > 0 : tri >= 0);
            if (ok) {
                const crt_entity& e = s.entities[hit_ent];
                const float4* rp = reinterpret_cast<const float4*>(rays + slot);
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
        float4* o = reinterpret_cast<float4*>(hits + slot);
        o[0] = tuvw;
        o[1] = make_float4(__int_as_float(tri), __int_as_float(mesh), __int_as_float(tri >= 0 ? hit_ent : -1),
                           __uint_as_float(pack_status(visits, status)));
        if (tri_tests) tri_tests[slot] = tests;
    }

    // Lane takes ray i.
    __device__ __forceinline__ void begin_ray(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                              crt_hit* __restrict__ hits, int64_t i, uint32_t flags,
                                              uint32_t* __restrict__ tri_tests) {
        active = true;
        leafing = false;
        walk_done = false;
        slot = int32_t(i);
        ray_tmax = reinterpret_cast<const float4*>(rays + i)[0].w;
        tmax = ray_tmax;
        ent = -1;
        hit_ent = -1;
        tri = mesh = -1;
        tests = status = visits = 0;
        if (!begin_walk(s, rays, flags)) {
            finish(s, rays, hits, tri_tests, flags);
            active = false;
        }
    }

    // Loop top (guards) and one entry.  kLeaves: leaves to test first, then
    // the walk goes on at cur, or ends (end_bits).
    __device__ __forceinline__ int32_t step(const SceneArgs& s, int32_t max_iter, NodeStack<LDS_DEPTH>& st) {
        if (it >= max_iter) {
            status |= CRT_STATUS_ITER_CAP;
            return kEnd;
        }
        if (cur < start || cur > start + count || cur < 0 || int64_t(cur) >= s.total_nodes) {
            status |= CRT_STATUS_BAD_INDEX;
            return kEnd;
        }
        it++;
        const float4* np = reinterpret_cast<const float4*>(s.nodes) + 4 * int64_t(cur);
        const float4 lmin = np[0], lmax = np[1], rmin = np[2], rmax = np[3];
        const int32_t lp = __float_as_int(lmin.w), rp = __float_as_int(rmin.w);
        const bool lleaf = lp != -1, rleaf = rp != -1;
        const float lt = lleaf ? -1.0f : bounds_stack(lmin, lmax, ro, inv, tmax);
        const float rt = rleaf ? -1.0f : bounds_stack(rmin, rmax, ro, inv, tmax);
        leaf0 = lp;
        leaf1 = rp;
        uint32_t end = 0;
        if (lt > 0.0f && rt > 0.0f) {
            int32_t next = __float_as_int(lmax.w) + start;
            int32_t post = __float_as_int(rmax.w) + start;
            if (rt < lt) {
                const int32_t tmp = next;
                next = post;
                post = tmp;
            }
            if (sp >= kStackCap) {
                end = kEndFlag | CRT_STATUS_STACK_OVERFLOW;
            } else {
                st.put(sp, post);
                sp++;
                cur = next;
            }
        } else if (lt > 0.0f) {
            cur = __float_as_int(lmax.w) + start;
        } else if (rt > 0.0f) {
            cur = __float_as_int(rmax.w) + start;
        } else if (sp <= 0) {
            end = kEndFlag;
        } else {
            --sp;
            cur = st.get(sp);
        }
        end_bits = end;
        if (leaf_nonempty(lp) || leaf_nonempty(rp)) return kLeaves;
        if (end) {
            status |= end & ~kEndFlag;
            return kEnd;
        }
        return kNext;
    }

    __device__ __forceinline__ bool test(int32_t idx, int4 tr, f3 a, f3 b, f3 c) {
        const float t = tri_t(ro, rd, a, b, c);
        tests++;
        if (t > 0.0f && t < tmax) {
            tmax = t;
            mesh = tr.w;
            tri = idx;
            hit_ent = ent;
            if (ANY) return true;
        }
        return false;
    }

    // One round of the wave: node phase, leaf phase, walk ends.
    __device__ __forceinline__ void round(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                          crt_hit* __restrict__ hits, uint32_t flags, int32_t max_iter,
                                          uint32_t* __restrict__ tri_tests, NodeStack<LDS_DEPTH>& st) {
        // --- node phase ----------------------------------------------------------
        for (int k = 0; k < INNER; ++k) {
            if (active && !walk_done && !leafing) {
                const int32_t r = step(s, max_iter, st);
                if (r == kEnd) {
                    walk_done = true;
                } else if (r == kLeaves) {
                    leafing = true;
                    const bool l0 = leaf_nonempty(leaf0);
                    leaf = l0 ? leaf0 : leaf1;
                    pi = l0 && leaf_nonempty(leaf1) ? 1 : 2;
                    off = 0;
                }
            }
            const unsigned long long stepping = __ballot(active && !walk_done && !leafing);
            if (stepping == 0 || __popcll(__ballot(leafing)) >= LEAF_BATCH) break;
        }
        // --- leaf phase: pending triangles, two per fetch ------------------------
        while (__ballot(leafing)) {
            if (leafing) {
                const int32_t len = leaf & 0xF;
                const int32_t first = (leaf >> 4) + off, left = len - off;
                // the second triangle, or the first again when one is left
                const int32_t second = first + (left > 1 ? 1 : 0);
                const int4 t0 = *reinterpret_cast<const int4*>(s.tris[first].packed);
                const int4 t1 = *reinterpret_cast<const int4*>(s.tris[second].packed);
                const f3 a0 = vertex_pos(s.verts, t0.x), b0 = vertex_pos(s.verts, t0.y), c0 = vertex_pos(s.verts, t0.z);
                const f3 a1 = vertex_pos(s.verts, t1.x), b1 = vertex_pos(s.verts, t1.y), c1 = vertex_pos(s.verts, t1.z);
                bool done = test(first, t0, a0, b0, c0);
                if (!done && left > 1) done = test(second, t1, a1, b1, c1);
                off += left > 1 ? 2 : 1;
                if (done) {   // an any-hit walk ends at its first accepted triangle
                    leafing = false;
                    walk_done = true;
                } else if (off >= len) {
                    if (pi == 1) {
                        leaf = leaf1;
                        pi = 2;
                        off = 0;
                    } else {
                        leafing = false;
                        if (end_bits) {
                            status |= end_bits & ~kEndFlag;
                            walk_done = true;
                        }
                    }
                }
            }
        }
        // --- walk ends: next entity or the ray's result --------------------------
        // the ends run once kWalkEndBatch lanes wait or no lane can step
        const unsigned long long ends = __ballot(walk_done);
        if (ends && (__popcll(ends) >= kWalkEndBatch || !__ballot(active && !walk_done && !leafing))) {
            if (walk_done) {
                walk_done = false;
                visits += uint32_t(it);
                bool fin = ANY && tri >= 0;   // any hit: the first entity with a hit ends the scene loop
                if (!fin) fin = !begin_walk(s, rays, flags);
                if (fin) {
                    finish(s, rays, hits, tri_tests, flags);
                    active = false;
                }
            }
        }
    }

    // Ray i per lane (i < 0: idle) until every lane is done.
    __device__ __forceinline__ void run(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                        crt_hit* __restrict__ hits, int64_t i, uint32_t flags, int32_t max_iter,
                                        uint32_t* __restrict__ tri_tests, NodeStack<LDS_DEPTH> st) {
        active = false;
        leafing = false;
        walk_done = false;
        if (i >= 0) begin_ray(s, rays, hits, i, flags, tri_tests);
        while (__ballot(active)) round(s, rays, hits, flags, max_iter, tri_tests, st);
    }

    // Rays [0, n) claimed from the sharded queue (queue.h): between rounds,
    // once at least `refill` lanes are idle, one claim takes a ray for each
    // idle lane.  Ends when the queue is dry and every lane is done.
    __device__ __forceinline__ void run_refill(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                               crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                               int32_t max_iter, uint32_t* __restrict__ tri_tests,
                                               NodeStack<LDS_DEPTH> st, unsigned long long* __restrict__ counters,
                                               int32_t refill) {
        active = false;
        leafing = false;
        walk_done = false;
        ShardQueue q;
        q.init(counters, n);
        while (true) {
            if (q.more) {
                const unsigned long long idle = __ballot(!active);
                const int32_t cnt = __popcll(idle);
                if (cnt >= refill) {
                    int32_t got = 0;
                    const int64_t first = q.claim(cnt, got);
                    if (!active) {
                        const int32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(idle >> 32),
                                                                       __builtin_amdgcn_mbcnt_lo(uint32_t(idle), 0));
                        if (rank < got) begin_ray(s, rays, hits, q.at(first + rank), flags, tri_tests);
                    }
                }
            }
            if (!__ballot(active)) {
                if (!q.more) break;
                continue;
            }
            round(s, rays, hits, flags, max_iter, tri_tests, st);
        }
    }
};

}  // namespace sdev
}  // namespace crt
