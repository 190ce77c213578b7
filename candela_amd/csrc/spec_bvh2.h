// Speculative wave walks of the reference's two BVH2 layouts, one template:
//
//   D == 0  the stackless layout (IntersectBVHStackless /
//           IntersectBVHStacklessOcclusion,
//           Source/Core/Shaders/Intersectors/Include/TraverseBVHStackless.glsl:175-278,
//           :463-556, in the IntersectScene entity loop :280-366, :558-579);
//   D > 0   the stack layout (IntersectBVHStack / IntersectBVHStackOcclusion,
//           Include/TraverseBVHStack.glsl:168-324, :509-657) with the first D
//           stack entries in LDS and the rest in per-lane scratch.
//
// Both reference walks decide the next node from box distances alone: the
// stackless walk follows a node's miss link after testing a leaf, whatever
// the test found, and the stack walk tests a node's leaf children and then
// goes on by the inner children's box distances.  A leaf's result only
// shortens TMax for the boxes that follow.  The wave walks of wave_walk.h and
// stack_walk.h park a lane at its leaves until enough lanes wait.  Here, as in
// the speculative wide walk (spec_walk.h):
//
//   node step  the reference's step (iteration cap, index guard, RayBounds
//              with the current TMax, then the next node), except that leaves
//              go to the lane's queue (Q slots in its LDS column) and the lane
//              goes straight on: a hit leaf of the stackless walk, every
//              non-empty leaf child (left, then right) of the stack walk.  A
//              lane steps while its queue has room for one more node's leaves.
//              A stack walk whose push finds 63 entries ends with
//              CRT_STATUS_STACK_OVERFLOW after its node's leaves are queued
//   round      up to INNER node steps (fewer once no lane can step), the
//              leaf phase whenever some queue holds a leaf, then the walk ends
//              once kWalkEndBatch lanes wait or no lane can step
//   leaf phase the queued leaves of the wave become one list (lane order, then
//              queue order); entry i is tested by lane i % 64 with the owner's
//              ray, one triangle per pass of the leaf; an accepted test
//              (0 < t < the owner's TMax) enters a 64-bit LDS atomic minimum per
//              owner over (t, order) for closest hits and (order, t) for any
//              hits, order = entry * 16 + triangle of the leaf
//   culling    boxes are tested with the TMax of the last leaf phase, which
//              only errs toward visiting more nodes; box distances, which
//              order the stack walk, do not depend on TMax, so the reference
//              walk's node sequence only gains whole subtrees and a
//              closest-hit walk still ends with the reference walk's nearest t
//
// As in spec_walk.h, a ray's record does not depend on which rays share its
// wave.  The host twin (oracle_trace.c: oracle_trace_stackless_spec,
// oracle_trace_stack_spec) replays the grid launch; run with a refilling
// preset's (INNER, Q) it gives that launch's records, counters included.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"
#include "device_math.h"
#include "queue.h"
#include "spec_walk.h"
#include "stack_walk.h"
#include "trace_launch.h"

namespace crt {
namespace sbvh2 {

using namespace ::crt::dm;
using spec::lanes_below;
using spec::wave_lds_sync;
using spec::WaveLds;

// ANY: the query; DYN: the query is the run-time value dyn_any instead (one
// walk for both queries, as the batched tiles need); COLS: the LDS column
// stride, the threads of the workgroup.
template <bool ANY, int D, int INNER, int Q, bool DYN = false, int COLS = 256>
struct SpecBvh2Wave {
    static_assert(Q >= 2 && Q <= 16, "queue slots fit 4 bits");
    static_assert(D >= 0 && D <= sdev::kStackCap, "LDS stack entries");
    static constexpr bool kStack = D > 0;
    static constexpr int32_t kPerNode = kStack ? 2 : 1;   // leaves one node step can queue

    // per lane
    f3 ro, rd, inv;
    float tmax, ray_tmax;
    int32_t start, count, ptr, sp, it;
    int32_t tri, mesh, slot, ent, hit_ent;
    uint32_t tests, status, visits;
    int32_t nq;          // queued leaves (this lane's LDS column, stride COLS)
    bool active, walk_done;
    bool dyn_any;        // the query when DYN (wave-uniform)
    LdsInt* q;
    sdev::NodeStack<kStack ? D : 1> st;   // the stack walk's node stack

    __device__ __forceinline__ bool any() const { return DYN ? dyn_any : ANY; }

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
        if (any()) tmax = ray_tmax;
        ro = xform(en.inverse, mk(a.x, a.y, a.z), 1.0f);
        rd = xform(en.inverse, mk(b.x, b.y, b.z), 0.0f);
        inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
        start = en.node_offset;
        count = en.node_count;
        ptr = start;
        sp = 0;
        it = 0;
        nq = 0;
        // the stackless `while (Pointer >= 0 ...)` never runs from a negative
        // root; the stack walk's index guard stops it instead
        walk_done = !kStack && ptr < 0;
        return true;
    }

    __device__ __forceinline__ void finish(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                           crt_hit* __restrict__ hits, uint32_t* __restrict__ tri_tests,
                                           uint32_t flags) {
        float4 tuvw = make_float4(-1.0f, -1.0f, -1.0f, -1.0f);
        const float closest = tri >= 0 ? tmax : -1.0f;
        if (any()) {
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

    __device__ __forceinline__ void begin_ray(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                              crt_hit* __restrict__ hits, int64_t i, uint32_t flags,
                                              uint32_t* __restrict__ tri_tests) {
        active = true;
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

    __device__ __forceinline__ bool can_step() const { return active && !walk_done && nq <= Q - kPerNode; }

    // One entry of the stack layout (TraverseBVHStack.glsl:168-324): both
    // children in 64 bytes, inner children box-tested, leaf children queued
    // without a box test.  False when the walk is over.
    __device__ __forceinline__ bool step_stack(const SceneArgs& s, int32_t max_iter) {
        if (it >= max_iter) {
            status |= CRT_STATUS_ITER_CAP;
            return false;
        }
        if (ptr < start || ptr > start + count || ptr < 0 || int64_t(ptr) >= s.total_nodes) {
            status |= CRT_STATUS_BAD_INDEX;
            return false;
        }
        it++;
        const float4* np = reinterpret_cast<const float4*>(s.nodes) + 4 * int64_t(ptr);
        const float4 lmin = np[0], lmax = np[1], rmin = np[2], rmax = np[3];
        const int32_t lp = __float_as_int(lmin.w), rp = __float_as_int(rmin.w);
        const float lt = lp != -1 ? -1.0f : bounds_stack(lmin, lmax, ro, inv, tmax);
        const float rt = rp != -1 ? -1.0f : bounds_stack(rmin, rmax, ro, inv, tmax);
        // slots nq and nq + 1 exist (can_step); each is kept only for a leaf
        // with triangles
        q[nq * COLS] = lp;
        nq += sdev::leaf_nonempty(lp) ? 1 : 0;
        q[nq * COLS] = rp;
        nq += sdev::leaf_nonempty(rp) ? 1 : 0;
        if (lt > 0.0f && rt > 0.0f) {
            int32_t next = __float_as_int(lmax.w) + start;
            int32_t post = __float_as_int(rmax.w) + start;
            if (rt < lt) {
                const int32_t tmp = next;
                next = post;
                post = tmp;
            }
            if (sp >= sdev::kStackCap) {
                status |= CRT_STATUS_STACK_OVERFLOW;
                return false;
            }
            st.put(sp, post);
            sp++;
            ptr = next;
        } else if (lt > 0.0f) {
            ptr = __float_as_int(lmax.w) + start;
        } else if (rt > 0.0f) {
            ptr = __float_as_int(rmax.w) + start;
        } else if (sp <= 0) {
            return false;
        } else {
            --sp;
            ptr = st.get(sp);
        }
        return true;
    }

    // One node of the stackless walk.  False when the walk is over.
    __device__ __forceinline__ bool step(const SceneArgs& s, int32_t max_iter) {
        if constexpr (kStack) return step_stack(s, max_iter);
        if (it >= max_iter) {
            status |= CRT_STATUS_ITER_CAP;
            return false;
        }
        if (ptr < start || ptr > start + count || int64_t(ptr) >= s.total_nodes) {
            status |= CRT_STATUS_BAD_INDEX;
            return false;
        }
        it++;
        const float4* __restrict__ nodes = reinterpret_cast<const float4*>(s.nodes);
        const float4 nmin = nodes[2 * int64_t(ptr)], nmax = nodes[2 * int64_t(ptr) + 1];
        const float tb = bounds_stackless(nmin, nmax, ro, inv, tmax);
        const int32_t link = __float_as_int(nmax.w);
        const int32_t packed = __float_as_int(nmin.w);
        const bool hit = tb > 0.0f && tb < tmax;
        if (hit && packed == -1) {
            ptr++;
            return true;
        }
        // the next free queue slot is written whatever the test says; a hit
        // leaf advances the count
        q[nq * COLS] = packed;
        nq += hit ? 1 : 0;
        if (link < 0) return false;
        ptr = link + start;
        return true;
    }

    // Tests every queued leaf of the wave, 64 at a time, and gives each owner
    // its winner; the queues are empty afterwards.
    __device__ __forceinline__ void leaf_phase(const SceneArgs& s, WaveLds<Q>& L) {
        const int32_t lane = __lane_id();
        // exclusive scan of the queue lengths (at most 16) by bit planes
        int32_t excl = 0, total = 0;
#pragma unroll
        for (int b = 0; b < 5; ++b) {
            const unsigned long long plane = __ballot((nq >> b) & 1);
            excl += lanes_below(plane) << b;
            total += __popcll(plane) << b;
        }
        for (int32_t j = 0; j < nq; ++j) {
            L.list[excl + j] = uint16_t((lane << 4) | j);
            tests += uint32_t(q[j * COLS] & 0xF);
        }
        L.keys[lane] = ~0ull;
        wave_lds_sync();
        for (int32_t base = 0; base < total; base += 64) {
            const int32_t idx = base + lane;
            const bool valid = idx < total;
            const uint32_t e = valid ? uint32_t(L.list[idx]) : 0u;
            const int32_t owner = int32_t(e >> 4), j = int32_t(e & 15u);
            // the owner's queue slot j sits (owner - lane) columns from this lane's
            const int32_t leaf = valid ? q[(owner - lane) + j * COLS] : 0;
            const float ox = __shfl(ro.x, owner), oy = __shfl(ro.y, owner), oz = __shfl(ro.z, owner);
            const float dx = __shfl(rd.x, owner), dy = __shfl(rd.y, owner), dz = __shfl(rd.z, owner);
            const float otmax = __shfl(tmax, owner);
            const int32_t first = leaf >> 4, len = leaf & 0xF;
            const f3 o = mk(ox, oy, oz), d = mk(dx, dy, dz);
            for (int32_t x = 0; __ballot(x < len); ++x) {
                const bool v = x < len;
                // a lane past its leaf reads triangle 0, which exists whenever a leaf does
                const int4 tr = *reinterpret_cast<const int4*>(s.tris[v ? first + x : 0].packed);
                const float t = tri_t(o, d, vertex_pos(s.verts, tr.x), vertex_pos(s.verts, tr.y),
                                      vertex_pos(s.verts, tr.z));
                if (v && t > 0.0f && t < otmax) {
                    const unsigned long long order = (unsigned long long)(uint32_t(idx) << 4 | uint32_t(x));
                    const unsigned long long tbits = __float_as_uint(t);
                    atomicMin(&L.keys[owner], any() ? ((order << 32) | tbits) : ((tbits << 32) | order));
                }
            }
        }
        wave_lds_sync();
        const unsigned long long k = L.keys[lane];
        if (k != ~0ull) {
            const uint32_t order = any() ? uint32_t(k >> 32) : uint32_t(k);
            const uint32_t tb = any() ? uint32_t(k) : uint32_t(k >> 32);
            const int32_t j = int32_t(L.list[order >> 4] & 15u);
            const int32_t idx = (q[j * COLS] >> 4) + int32_t(order & 15u);
            tmax = __uint_as_float(tb);
            tri = idx;
            mesh = s.tris[idx].packed[3];
            hit_ent = ent;
            if (any()) walk_done = true;   // the first accepted triangle ends an any-hit walk
        }
        nq = 0;
        wave_lds_sync();   // the list and keys are reused by the next phase
    }

    __device__ __forceinline__ void round(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                          crt_hit* __restrict__ hits, uint32_t flags, int32_t max_iter,
                                          uint32_t* __restrict__ tri_tests, WaveLds<Q>& L) {
        for (int k = 0; k < INNER; ++k) {
            if (can_step()) {
                if (!step(s, max_iter)) walk_done = true;
            }
            if (!__ballot(can_step())) break;
        }
        if (__ballot(nq > 0)) leaf_phase(s, L);
        // every queue is empty here; the ends run once kWalkEndBatch lanes
        // wait or no lane can step
        const bool ends = active && walk_done;
        const unsigned long long em = __ballot(ends);
        if (em && (__popcll(em) >= kWalkEndBatch || !__ballot(can_step()))) {
            if (ends) {
                visits += uint32_t(it);
                bool fin = any() && tri >= 0;
                if (!fin) fin = !begin_walk(s, rays, flags);
                if (fin) {
                    finish(s, rays, hits, tri_tests, flags);
                    active = false;
                }
            }
        }
    }

    // Ray i per lane (i < 0: lane idle) until every lane is done.
    __device__ __forceinline__ void run(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                        crt_hit* __restrict__ hits, int64_t i, uint32_t flags, int32_t max_iter,
                                        uint32_t* __restrict__ tri_tests, WaveLds<Q>& L) {
        active = false;
        walk_done = false;
        nq = 0;
        if (i >= 0) begin_ray(s, rays, hits, i, flags, tri_tests);
        while (__ballot(active)) round(s, rays, hits, flags, max_iter, tri_tests, L);
    }

    // Rays [0, n) claimed from the sharded queue between rounds, one per idle
    // lane, once at least `refill` lanes are idle.
    __device__ __forceinline__ void run_refill(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                               crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                               int32_t max_iter, uint32_t* __restrict__ tri_tests, WaveLds<Q>& L,
                                               unsigned long long* __restrict__ counters, int32_t refill) {
        active = false;
        walk_done = false;
        nq = 0;
        ShardQueue sq;
        sq.init(counters, n);
        while (true) {
            if (sq.more) {
                const unsigned long long idle = __ballot(!active);
                const int32_t cnt = __popcll(idle);
                if (cnt >= refill) {
                    int32_t got = 0;
                    const int64_t first = sq.claim(cnt, got);
                    if (!active) {
                        const int32_t rank = lanes_below(idle);
                        if (rank < got) begin_ray(s, rays, hits, sq.at(first + rank), flags, tri_tests);
                    }
                }
            }
            if (!__ballot(active)) {
                if (!sq.more) break;
                continue;
            }
            round(s, rays, hits, flags, max_iter, tri_tests, L);
        }
    }
};

}  // namespace sbvh2
}  // namespace crt
