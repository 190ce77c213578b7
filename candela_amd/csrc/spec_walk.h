// Speculative wave walk of the wide layout (BVH4 / BVH8): lanes never wait
// for their leaf tests, and the wave tests leaves 64 at a time.
//
// The wave-scheduled walk of wide_walk.h parks a lane whose node step found
// leaves until the wave's leaf phase, because the lane's next box tests need
// the tmax those tests leave.  On incoherent rays that idles most lanes of a
// node step, and the leaf phase itself runs one lane per ray with few rays
// waiting.  Here:
//
//   node step  as wide_walk.h's (same slab arithmetic, exit distance and tmax
//              grown by kWideGrow, children sorted nearest first), but every
//              hit inner child except the nearest is pushed and the lane goes
//              straight on to the nearest, whether or not the node had leaves;
//              the hit leaves go to the lane's own queue (Q slots in its LDS
//              column, no work across lanes); a lane steps only while its
//              queue has room for a whole node's leaves
//   round      up to INNER node steps (fewer once no lane can step), then the
//              leaf phase whenever some queue holds a leaf, then the walk ends
//              (next entity, or the ray's record) of the lanes whose walks are
//              over, once kEndBatch of them wait or no lane can step: the end
//              code costs the whole wave, so it runs for many lanes at once
//   leaf phase the queued leaves of all lanes become one list: lane order,
//              then queue order (an exclusive scan of the queue lengths by
//              ballot bit planes, then each lane writes (lane, slot) entries);
//              list entry i is tested by lane i % 64 with the owner's ray
//              (lane shuffles), the triangles of a leaf two per fetch; an
//              accepted test (0 < t < the owner's tmax) enters a 64-bit LDS
//              atomic minimum per owner over (t, order) for closest hits and
//              (order, t) for any hits, where order = entry * 16 + triangle
//              of the leaf; then each owner takes its winner (closest: nearer
//              hit; any: the walk ends) and every queue is empty
//   culling    box tests and pops use the tmax of the last leaf phase, which
//              only errs toward visiting more nodes, so a closest-hit walk
//              still reports the nearest triangle that any wide walk finds
//
// A ray's record does not depend on which rays share its wave.  In a round a
// lane that can step always steps, and a lane that cannot (queue full, walk
// over) cannot become able again before the leaf phase; so every lane takes
// INNER node steps, or fewer when its own queue fills or its walk ends, and
// then its own leaves are tested, whatever the other lanes do.  Walk ends and
// refills happen between rounds and only delay a lane.  The host twin
// (oracle_trace.c: oracle_trace_wide_spec) replays the grid launch; run with a
// refilling preset's (LDS depth, INNER, Q) it gives that launch's records too,
// counters included, whichever rays the waves claimed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"
#include "device_math.h"
#include "queue.h"
#include "trace_launch.h"
#include "wide_walk.h"

namespace crt {
namespace spec {

using namespace ::crt::dm;
using wide::kWideGrow;
using wide::Stack;

// Ints per lane column: the stack's 2 * LDS_DEPTH, then the Q queue slots.
template <int LDS_DEPTH, int Q>
constexpr int lds_ints() {
    return 2 * LDS_DEPTH + Q;
}

// LDS of one wave: the leaf-phase list ((lane << 4) | slot per queued leaf)
// and the 64 owners' keys.
template <int Q>
struct WaveLds {
    uint16_t list[64 * Q];
    unsigned long long keys[64];
};

__device__ __forceinline__ int32_t lanes_below(unsigned long long mask) {
    return int32_t(__builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u)));
}

// Orders the LDS accesses of the lanes of one wave (list and queue entries
// written by one lane and read by another) without a workgroup barrier.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// A queued leaf packs its first record above 4 count bits.
constexpr int32_t kMaxRecord = 1 << 26;

// Lanes whose walks are over wait for the walk-end code until this many of
// them do, or until no lane can step (measured on config 2's GI rays: 32
// takes the refilling launch from 0.590 to 0.555 ms; 16 gives 0.585, 48 0.562).
constexpr int32_t kEndBatch = kWalkEndBatch;

// The query is ANY, or with DYN the member dyn_any, set per ray batch (one
// compiled walk for kernels that serve both queries: the batched kernel).
template <bool ANY, int W, int LDS_DEPTH, int INNER, int Q, bool DYN = false>
struct SpecWave {
    static constexpr int G = W / 4;
    static constexpr int32_t kCap = LDS_DEPTH + wide::kSpill;
    static_assert(Q >= W && Q <= 16, "a queue holds one node's leaves; slots fit 4 bits");

    bool dyn_any;
    __device__ __forceinline__ bool any() const { return DYN ? dyn_any : ANY; }

    // per lane
    f3 ro, rd, inv;
    float tmax, ray_tmax;
    int32_t root, cur, sp, it;
    uint32_t span;       // node indices root .. root + span - 1 pass the index guard
    int32_t tri, mesh, slot, ent, hit_ent;
    uint32_t tests, status, visits;
    int32_t nq;          // queued leaves (in the lane's LDS column)
    bool active, walk_done;

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
        begin_local(xform(en.inverse, mk(a.x, a.y, a.z), 1.0f), xform(en.inverse, mk(b.x, b.y, b.z), 0.0f),
                    en.node_offset, en.node_count, s.total_nodes);
        return true;
    }

    // A walk of one entity's tree from the ray (lro, lrd) in its local space,
    // with the current tmax: empty stack and queue, the index guard of
    // wide_walk.h's Core::begin.  (The path kernel starts its segments here.)
    __device__ __forceinline__ void begin_local(f3 lro, f3 lrd, int32_t r, int32_t c, int64_t total_nodes) {
        ro = lro;
        rd = lrd;
        inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
        root = r;
        const int64_t end = int32_t(uint32_t(r) + uint32_t(c));
        const int64_t lim = end < total_nodes ? end : total_nodes;
        span = lim > r ? uint32_t(lim - r) : 0u;
        cur = r;
        sp = 0;
        it = 0;
        nq = 0;
        walk_done = false;
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
                const float4* r = reinterpret_cast<const float4*>(s.tri_records) + 3 * int64_t(tri);
                const float4 va = r[0], vb = r[1], vc = r[2];
                const f3 p = mk(lro.x + lrd.x * closest, lro.y + lrd.y * closest, lro.z + lrd.z * closest);
                const f3 uvw = barycentrics(p, mk(va.x, va.y, va.z), mk(vb.x, vb.y, vb.z), mk(vc.x, vc.y, vc.z));
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

    // A lane steps while its walk goes on and its queue has room for W leaves.
    __device__ __forceinline__ bool can_step() const { return active && !walk_done && nq <= Q - W; }

    // Pop (culling by the current tmax), guards and one node; the node's hit
    // leaves are appended to the lane's queue.  False when the walk is over.
    __device__ __forceinline__ bool step(const SceneArgs& s, int32_t max_iter, Stack<LDS_DEPTH>& st) {
        if (cur < 0) {
            bool found = false;
            while (sp > 0) {
                --sp;
                int32_t node;
                float t;
                st.get(sp, node, t);
                if (!any() && t > tmax * kWideGrow) continue;
                cur = node;
                found = true;
                break;
            }
            if (!found) return false;
        }
        if (it >= max_iter) {
            status |= CRT_STATUS_ITER_CAP;
            return false;
        }
        if (uint32_t(cur) - uint32_t(root) >= span) {
            status |= CRT_STATUS_BAD_INDEX;
            return false;
        }
        it++;
        int32_t nl = 0;   // leaves of this node, written at queue slots nq, nq + 1, ...
        int32_t m = 0;    // hit inner children with a node index >= 0
        const float4* __restrict__ groups = reinterpret_cast<const float4*>(s.nodes);
        const float inf = __builtin_inff();
        const float tgrow = tmax * kWideGrow;
        float t0 = inf, t1 = inf, t2 = inf, t3 = inf, t4 = inf, t5 = inf, t6 = inf, t7 = inf;
        int32_t n0 = -1, n1 = -1, n2 = -1, n3 = -1, n4 = -1, n5 = -1, n6 = -1, n7 = -1;
#define CRT_SPEC_CHILD(LO_X, HI_X, LO_Y, HI_Y, LO_Z, HI_Z, CH, CN, T, N)                                       \
    {                                                                                                       \
        const float fx = (HI_X - ro.x) * inv.x, fy = (HI_Y - ro.y) * inv.y, fz = (HI_Z - ro.z) * inv.z;     \
        const float nx = (LO_X - ro.x) * inv.x, ny = (LO_Y - ro.y) * inv.y, nz = (LO_Z - ro.z) * inv.z;     \
        const float e1 = mn(mn(mx(fx, nx), mn(mx(fy, ny), mx(fz, nz))) * kWideGrow, tgrow);                \
        const float e0 = mx(mx(mn(fx, nx), mx(mn(fy, ny), mn(fz, nz))), 0.0f);                              \
        const int32_t cnt = __float_as_int(CN), chi = __float_as_int(CH);                                   \
        const bool hit = cnt >= 0 && e1 >= e0 && e1 > 0.0f;                                                 \
        /* the next free queue slot is written whatever the test says; only a */                            \
        /* hit leaf whose records fit a slot's packing advances the count      */                            \
        const bool take = hit && cnt > 0;                                                                   \
        const bool fits = chi >= 0 && chi <= kMaxRecord - cnt;                                              \
        st.put_leaf(nq + nl, (chi << 4) | cnt);                                                             \
        nl += take && fits ? 1 : 0;                                                                         \
        status |= take && !fits ? uint32_t(CRT_STATUS_BAD_INDEX) : 0u;                                      \
        const bool inner = hit && cnt == 0;                                                                 \
        T = inner ? e0 : inf;                                                                               \
        N = inner ? chi : -1;                                                                               \
        m += N >= 0 ? 1 : 0;                                                                                \
    }
#define CRT_SPEC_GROUP(G_, TA, NA, TB, NB, TC, NC, TD, ND)                                                  \
    {                                                                                                       \
        const float4* p = groups + (int64_t(cur) * G + G_) * 8;                                             \
        const float4 lox = p[0], hix = p[1], loy = p[2], hiy = p[3], loz = p[4], hiz = p[5];                \
        const float4 chf = p[6], cnf = p[7];                                                                \
        CRT_SPEC_CHILD(lox.x, hix.x, loy.x, hiy.x, loz.x, hiz.x, chf.x, cnf.x, TA, NA)                      \
        CRT_SPEC_CHILD(lox.y, hix.y, loy.y, hiy.y, loz.y, hiz.y, chf.y, cnf.y, TB, NB)                      \
        CRT_SPEC_CHILD(lox.z, hix.z, loy.z, hiy.z, loz.z, hiz.z, chf.z, cnf.z, TC, NC)                      \
        CRT_SPEC_CHILD(lox.w, hix.w, loy.w, hiy.w, loz.w, hiz.w, chf.w, cnf.w, TD, ND)                      \
    }
#define CRT_SPEC_CX(TI, NI, TJ, NJ)         \
    {                                       \
        const bool sw = TJ < TI;            \
        const float ta = TI, tb = TJ;       \
        const int32_t na = NI, nb = NJ;     \
        TI = sw ? tb : ta;                  \
        TJ = sw ? ta : tb;                  \
        NI = sw ? nb : na;                  \
        NJ = sw ? na : nb;                  \
    }
        CRT_SPEC_GROUP(0, t0, n0, t1, n1, t2, n2, t3, n3)
        if constexpr (W == 8) CRT_SPEC_GROUP(1, t4, n4, t5, n5, t6, n6, t7, n7)
        const int32_t pushes = m > 0 ? m - 1 : 0;
        if (sp + pushes > kCap) {   // the node's leaves are dropped with the walk
            status |= CRT_STATUS_STACK_OVERFLOW;
            return false;
        }
        if constexpr (W == 4) {
            CRT_SPEC_CX(t0, n0, t1, n1) CRT_SPEC_CX(t2, n2, t3, n3) CRT_SPEC_CX(t0, n0, t2, n2)
            CRT_SPEC_CX(t1, n1, t3, n3) CRT_SPEC_CX(t1, n1, t2, n2)
        } else {
            CRT_SPEC_CX(t0, n0, t1, n1) CRT_SPEC_CX(t2, n2, t3, n3) CRT_SPEC_CX(t4, n4, t5, n5)
            CRT_SPEC_CX(t6, n6, t7, n7) CRT_SPEC_CX(t0, n0, t2, n2) CRT_SPEC_CX(t1, n1, t3, n3)
            CRT_SPEC_CX(t4, n4, t6, n6) CRT_SPEC_CX(t5, n5, t7, n7) CRT_SPEC_CX(t1, n1, t2, n2)
            CRT_SPEC_CX(t5, n5, t6, n6) CRT_SPEC_CX(t0, n0, t4, n4) CRT_SPEC_CX(t1, n1, t5, n5)
            CRT_SPEC_CX(t2, n2, t6, n6) CRT_SPEC_CX(t3, n3, t7, n7) CRT_SPEC_CX(t2, n2, t4, n4)
            CRT_SPEC_CX(t3, n3, t5, n5) CRT_SPEC_CX(t1, n1, t2, n2) CRT_SPEC_CX(t3, n3, t4, n4)
            CRT_SPEC_CX(t5, n5, t6, n6)
        }
#undef CRT_SPEC_CX
#undef CRT_SPEC_GROUP
#undef CRT_SPEC_CHILD
        // farthest first, so that pops come nearest first; the nearest is next.
        // Sorted entry k (1 <= k < m) goes to stack index sp + m - 1 - k.
        if (sp + m <= LDS_DEPTH) {
            // the run stays in LDS: entry k on row LDS_DEPTH - sp - m + k, all from one base
            LdsInt* b = st.row(LDS_DEPTH - sp - m);
            const int32_t rs = 2 * st.stride;
#define CRT_SPEC_PUT(K, N, T)                           \
    if (K < m) {                                        \
        b[K * rs] = N;                                  \
        b[K * rs + st.stride] = __float_as_int(T);      \
    }
            if constexpr (W == 8) {
                CRT_SPEC_PUT(7, n7, t7)
                CRT_SPEC_PUT(6, n6, t6)
                CRT_SPEC_PUT(5, n5, t5)
                CRT_SPEC_PUT(4, n4, t4)
            }
            CRT_SPEC_PUT(3, n3, t3)
            CRT_SPEC_PUT(2, n2, t2)
            CRT_SPEC_PUT(1, n1, t1)
#undef CRT_SPEC_PUT
            sp += pushes;
        } else {
            if constexpr (W == 8) {
                if (7 < m) st.put(sp++, n7, t7);
                if (6 < m) st.put(sp++, n6, t6);
                if (5 < m) st.put(sp++, n5, t5);
                if (4 < m) st.put(sp++, n4, t4);
            }
            if (3 < m) st.put(sp++, n3, t3);
            if (2 < m) st.put(sp++, n2, t2);
            if (1 < m) st.put(sp++, n1, t1);
        }
        cur = n0;   // -1 (no inner child hit): pop first
        nq += nl;
        return true;
    }

    // Tests every queued leaf of the wave, 64 at a time, and gives each owner
    // its winner; the queues are empty afterwards.
    __device__ __forceinline__ void leaf_phase(const SceneArgs& s, Stack<LDS_DEPTH>& st, WaveLds<Q>& L) {
        const int32_t lane = __lane_id();
        const float4* __restrict__ rec = reinterpret_cast<const float4*>(s.tri_records);
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
            tests += uint32_t(st.leaf(j) & 0xF);
        }
        L.keys[lane] = ~0ull;
        wave_lds_sync();
        // the owner's queue slot j sits (owner - lane) columns from this lane's
        const LdsInt* col = st.col();
        const int32_t qrow = 2 * LDS_DEPTH * st.stride;
        for (int32_t base = 0; base < total; base += 64) {
            const int32_t idx = base + lane;
            const bool valid = idx < total;
            const uint32_t e = valid ? uint32_t(L.list[idx]) : 0u;
            const int32_t owner = int32_t(e >> 4), j = int32_t(e & 15u);
            const int32_t leaf = valid ? col[(owner - lane) + qrow + j * st.stride] : 0;
            const float ox = __shfl(ro.x, owner), oy = __shfl(ro.y, owner), oz = __shfl(ro.z, owner);
            const float dx = __shfl(rd.x, owner), dy = __shfl(rd.y, owner), dz = __shfl(rd.z, owner);
            const float otmax = __shfl(tmax, owner);
            const int32_t first = leaf >> 4, len = leaf & 0xF;
            const f3 o = mk(ox, oy, oz), d = mk(dx, dy, dz);
            for (int32_t x = 0; __ballot(x < len); x += 2) {
                const bool v0 = x < len, v1 = x + 1 < len;
                // a lane past its leaf reads record 0, which exists whenever a leaf does
                const float4* p = rec + 3 * int64_t(v0 ? first + x : 0);
                const float4* p1 = p + (v1 ? 3 : 0);
                const float4 a0 = p[0], b0 = p[1], c0 = p[2];
                const float4 a1 = p1[0], b1 = p1[1], c1 = p1[2];
                const float ta = tri_t(o, d, mk(a0.x, a0.y, a0.z), mk(b0.x, b0.y, b0.z), mk(c0.x, c0.y, c0.z));
                const float tb = tri_t(o, d, mk(a1.x, a1.y, a1.z), mk(b1.x, b1.y, b1.z), mk(c1.x, c1.y, c1.z));
                const unsigned long long order = (unsigned long long)(uint32_t(idx) << 4 | uint32_t(x));
                if (v0 && ta > 0.0f && ta < otmax) {
                    const unsigned long long tbits = __float_as_uint(ta);
                    atomicMin(&L.keys[owner], any() ? ((order << 32) | tbits) : ((tbits << 32) | order));
                }
                if (v1 && tb > 0.0f && tb < otmax) {
                    const unsigned long long tbits = __float_as_uint(tb);
                    atomicMin(&L.keys[owner], any() ? (((order + 1) << 32) | tbits) : ((tbits << 32) | (order + 1)));
                }
            }
        }
        wave_lds_sync();
        const unsigned long long k = L.keys[lane];
        if (k != ~0ull) {
            const uint32_t order = any() ? uint32_t(k >> 32) : uint32_t(k);
            const uint32_t tb = any() ? uint32_t(k) : uint32_t(k >> 32);
            const int32_t j = int32_t(L.list[order >> 4] & 15u);
            const float4* p = rec + 3 * int64_t((st.leaf(j) >> 4) + int32_t(order & 15u));
            tmax = __uint_as_float(tb);
            tri = __float_as_int(p[0].w);
            mesh = __float_as_int(p[1].w);
            hit_ent = ent;
            if (any()) walk_done = true;   // the first accepted triangle ends an any-hit walk
        }
        nq = 0;
        wave_lds_sync();   // the list and keys are reused by the next phase
    }

    // One round of the wave: node steps, leaf phase, walk ends.
    __device__ __forceinline__ void round(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                          crt_hit* __restrict__ hits, uint32_t flags, int32_t max_iter,
                                          uint32_t* __restrict__ tri_tests, Stack<LDS_DEPTH>& st, WaveLds<Q>& L) {
        for (int k = 0; k < INNER; ++k) {
            if (can_step()) {
                if (!step(s, max_iter, st)) walk_done = true;
            }
            if (!__ballot(can_step())) break;
        }
        if (__ballot(nq > 0)) leaf_phase(s, st, L);
        // every queue is empty here: a walk ends once it has no step left, and
        // the ends run once kEndBatch lanes wait or no lane can step
        const bool ends = active && walk_done;
        const unsigned long long em = __ballot(ends);
        if (em && (__popcll(em) >= kEndBatch || !__ballot(can_step()))) {
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

    __device__ __forceinline__ void run(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                        crt_hit* __restrict__ hits, int64_t i, uint32_t flags, int32_t max_iter,
                                        uint32_t* __restrict__ tri_tests, Stack<LDS_DEPTH> st, WaveLds<Q>& L) {
        active = false;
        walk_done = false;
        nq = 0;
        if (i >= 0) begin_ray(s, rays, hits, i, flags, tri_tests);
        while (__ballot(active)) round(s, rays, hits, flags, max_iter, tri_tests, st, L);
    }

    // Rays [0, n) claimed from the sharded queue between rounds, one per idle
    // lane, once at least `refill` lanes are idle.
    __device__ __forceinline__ void run_refill(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                               crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                               int32_t max_iter, uint32_t* __restrict__ tri_tests,
                                               Stack<LDS_DEPTH> st, WaveLds<Q>& L,
                                               unsigned long long* __restrict__ counters, int32_t refill) {
        active = false;
        walk_done = false;
        nq = 0;
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
                        const int32_t rank = lanes_below(idle);
                        if (rank < got) begin_ray(s, rays, hits, q.at(first + rank), flags, tri_tests);
                    }
                }
            }
            if (!__ballot(active)) {
                if (!q.more) break;
                continue;
            }
            round(s, rays, hits, flags, max_iter, tri_tests, st, L);
        }
    }
};

}  // namespace spec
}  // namespace crt
