// Walks of the wide layout (CRT_FORMAT_WIDE4 / WIDE8, crt_wide_build).
//
// Per ray and entity the walk is one fixed sequence of node steps and leaf
// tests, so the one-ray-per-lane walk (walk_wide_thread, used by the scene
// loop of thread_walk.h), the wave-scheduled walk (WideWave, wide.hip) and the
// host twin (oracle_trace.c: walk_wide) agree bit for bit, node visits and
// triangle tests included:
//
//   pop       when no next node is known: pop (node, entry distance) entries
//             until one's entry distance is within the grown tmax (closest
//             hit; any hit pops without culling); an empty stack ends the walk
//   guards    iteration cap (CRT_STATUS_ITER_CAP), then node index within the
//             entity's [root, root + count) (CRT_STATUS_BAD_INDEX)
//   node      every non-empty child box gets the slab arithmetic of the stack
//             include's RayBounds (Include/TraverseBVHStack.glsl:107-116), with
//             the exit distance and tmax grown by kWideGrow; hit leaves are
//             appended to the pending list in slot order; hit inner children
//             are sorted by entry distance with a fixed compare-exchange
//             network and pushed farthest first, all but the nearest; any-hit
//             walks with CRT_FLAG_SLOT_ORDER skip the sort: the first hit child
//             in slot order comes first, the others are pushed last slot first
//   leaves    with pending leaves the first child is pushed too (on top),
//             the pending triangles are tested in slot order, and the next
//             node is popped (with the tmax those tests produced); without,
//             the walk goes to the first child, else pops
//   overflow  pushes that would pass LDS_DEPTH + 64 entries end the walk with
//             CRT_STATUS_STACK_OVERFLOW before any leaf of that node is tested
//
// Triangle tests and acceptance (0 < t < tmax, first found wins a tie) are
// those of test_leaf (TraverseBVHStackless.glsl:79-97, :222-235); records are
// three 16-byte loads (v0 + triangle index, v1 + mesh id, v2).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"
#include "device_math.h"
#include "queue.h"
#include "trace_launch.h"

namespace crt {
namespace wide {

using namespace ::crt::dm;

// Relative margin of the wide box tests.  A triangle lying on a face of its
// leaf box (touching boxes, axis-aligned geometry) can have a computed t a
// few ulps below the box's computed entry distance; with a closer-by-ulps hit
// already found, an exact comparison would cull the leaf holding the true
// nearest triangle.  Growing the exit distance and tmax by 2^-19 (16 ulps)
// keeps such leaves, so the walk reports the brute-force nearest hit.
constexpr float kWideGrow = 1.0f + 1.0f / 524288.0f;
constexpr int kSpill = 64;   // stack entries beyond the LDS part, in per-lane scratch

enum StepResult : int32_t { kNext = 0, kLeaves = 1, kEnd = 2 };

// Per-lane LDS column (stride `stride` ints, entry-major, so lane i of a wave
// touches bank i): 2 * LDS_DEPTH ints of stack entries (node, entry
// distance), then W ints of pending leaves.  Stack entries beyond LDS_DEPTH
// go to per-lane arrays of kSpill entries, separate locals of the caller (not
// members), so that the walk state around them can live in registers; the
// pending leaves sit in LDS for the same reason (a register array picked by a
// data-dependent index would be kept in scratch).
template <int W, int LDS_DEPTH>
constexpr int lds_ints() {
    return 2 * LDS_DEPTH + W;
}

template <int LDS_DEPTH>
struct Stack {
    int32_t* lds;
    int32_t stride;
    int32_t* spill_node;
    float* spill_t;

    __device__ __forceinline__ LdsInt* col() const { return (LdsInt*)lds; }   // see LdsInt
    // Stack entry sp < LDS_DEPTH sits on row LDS_DEPTH - 1 - sp (its node, and its entry distance one
    // column stride later): the stack grows toward row 0, so the entries one node step pushes lie on
    // consecutive rows above one base row and are stored at constant offsets from one address.
    __device__ __forceinline__ LdsInt* row(int32_t r) const { return col() + 2 * r * stride; }

    __device__ __forceinline__ void put_leaf(int32_t j, int32_t v) { col()[(2 * LDS_DEPTH + j) * stride] = v; }
    __device__ __forceinline__ int32_t leaf(int32_t j) const { return col()[(2 * LDS_DEPTH + j) * stride]; }

    __device__ __forceinline__ void put(int32_t sp, int32_t node, float t) {
        if (sp < LDS_DEPTH) {
            LdsInt* r = row(LDS_DEPTH - 1 - sp);
            r[0] = node;
            r[stride] = __float_as_int(t);
        } else {
            spill_node[sp - LDS_DEPTH] = node;
            spill_t[sp - LDS_DEPTH] = t;
        }
    }
    __device__ __forceinline__ void get(int32_t sp, int32_t& node, float& t) const {
        if (sp < LDS_DEPTH) {
            const LdsInt* r = row(LDS_DEPTH - 1 - sp);
            node = r[0];
            t = __int_as_float(r[stride]);
        } else {
            node = spill_node[sp - LDS_DEPTH];
            t = spill_t[sp - LDS_DEPTH];
        }
    }
};

// One ray's walk of one entity's wide tree.  The query is ANY, or with DYN
// the member dyn_any, set per walk (one compiled walk serving both queries,
// for kernels that mix them: two inlined walks would add up their registers).
template <bool ANY, int W, int LDS_DEPTH, bool DYN = false>
struct Core {
    static constexpr int G = W / 4;
    static constexpr int32_t kCap = LDS_DEPTH + kSpill;

    bool dyn_any;
    __device__ __forceinline__ bool any() const { return DYN ? dyn_any : ANY; }

    f3 ro, rd, inv;
    float tmax;          // closest: shrinks with every accepted hit
    int32_t root;
    uint32_t span;       // node indices root .. root + span - 1 pass the index guard
    int32_t cur;         // next node, or -1: pop first
    int32_t sp, it;
    int32_t np, pi;      // pending leaves (in LDS, start << 4 | count): count, next to test
    int32_t tri, mesh;   // accepted hit of this walk
    uint32_t tests, status;

    // The index guard passes cur in [r, min(r + c, total_nodes)), with r + c in 32-bit arithmetic as the
    // host walk computes it; span is that range's length (at most 2^31 - 1 - r), so the guard is one
    // unsigned compare of cur - r mod 2^32 per step: for cur < r that difference is at least 2^31 - r.
    __device__ __forceinline__ void begin(f3 lro, f3 lrd, float walk_tmax, int32_t r, int32_t c,
                                          int64_t total_nodes) {
        ro = lro;
        rd = lrd;
        inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
        tmax = walk_tmax;
        root = r;
        const int64_t end = int32_t(uint32_t(r) + uint32_t(c));
        const int64_t lim = end < total_nodes ? end : total_nodes;
        span = lim > r ? uint32_t(lim - r) : 0u;
        cur = r;
        sp = 0;
        it = 0;
        np = 0;
        pi = 0;
    }

    // Pop, guards, one node.  kNext: cur holds the next node; kLeaves: np
    // pending leaves to test before the next step; kEnd: the walk is over.
    // slot_order (any hit only, CRT_FLAG_SLOT_ORDER): hit children in slot
    // order instead of nearest first.
    __device__ __forceinline__ int32_t step(const SceneArgs& s, int32_t max_iter, Stack<LDS_DEPTH>& st,
                                            bool slot_order) {
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
            if (!found) return kEnd;
        }
        if (it >= max_iter) {
            status |= CRT_STATUS_ITER_CAP;
            return kEnd;
        }
        if (uint32_t(cur) - uint32_t(root) >= span) {
            status |= CRT_STATUS_BAD_INDEX;
            return kEnd;
        }
        it++;
        np = 0;
        pi = 0;
        int32_t m = 0;   // hit inner children with a node index >= 0 (the entries the sort puts first)
        const float4* __restrict__ groups = reinterpret_cast<const float4*>(s.nodes);
        // hit inner children as (entry distance, node) in named registers t0..t7 /
        // n0..n7 (no arrays: every index is a compile-time name); misses and
        // leaves are (+inf, -1)
        const float inf = __builtin_inff();
        const float tgrow = tmax * kWideGrow;
        float t0 = inf, t1 = inf, t2 = inf, t3 = inf, t4 = inf, t5 = inf, t6 = inf, t7 = inf;
        int32_t n0 = -1, n1 = -1, n2 = -1, n3 = -1, n4 = -1, n5 = -1, n6 = -1, n7 = -1;
#define CRT_WIDE_CHILD(LO_X, HI_X, LO_Y, HI_Y, LO_Z, HI_Z, CH, CN, T, N)                                       \
    {                                                                                                       \
        const float fx = (HI_X - ro.x) * inv.x, fy = (HI_Y - ro.y) * inv.y, fz = (HI_Z - ro.z) * inv.z;     \
        const float nx = (LO_X - ro.x) * inv.x, ny = (LO_Y - ro.y) * inv.y, nz = (LO_Z - ro.z) * inv.z;     \
        const float e1 = mn(mn(mx(fx, nx), mn(mx(fy, ny), mx(fz, nz))) * kWideGrow, tgrow);                \
        const float e0 = mx(mx(mn(fx, nx), mx(mn(fy, ny), mn(fz, nz))), 0.0f);                              \
        const int32_t cnt = __float_as_int(CN), chi = __float_as_int(CH);                                   \
        /* e0 >= 0, so (e0 > 0 || e1 > 0) given e1 >= e0 is e1 > 0 */                                      \
        const bool hit = cnt >= 0 && e1 >= e0 && e1 > 0.0f;                                                 \
        /* the next free pending slot is written whatever the test says; a */                               \
        /* hit leaf advances the count                                      */                               \
        st.put_leaf(np, (chi << 4) | cnt);                                                                  \
        np += hit && cnt > 0 ? 1 : 0;                                                                       \
        const bool inner = hit && cnt == 0;                                                                 \
        T = inner ? e0 : inf;                                                                               \
        N = inner ? chi : -1;                                                                               \
        m += N >= 0 ? 1 : 0;                                                                                \
    }
#define CRT_WIDE_GROUP(G_, TA, NA, TB, NB, TC, NC, TD, ND)                                                  \
    {                                                                                                       \
        const float4* p = groups + (int64_t(cur) * G + G_) * 8;                                             \
        const float4 lox = p[0], hix = p[1], loy = p[2], hiy = p[3], loz = p[4], hiz = p[5];                \
        const float4 chf = p[6], cnf = p[7];                                                                \
        CRT_WIDE_CHILD(lox.x, hix.x, loy.x, hiy.x, loz.x, hiz.x, chf.x, cnf.x, TA, NA)                      \
        CRT_WIDE_CHILD(lox.y, hix.y, loy.y, hiy.y, loz.y, hiz.y, chf.y, cnf.y, TB, NB)                      \
        CRT_WIDE_CHILD(lox.z, hix.z, loy.z, hiy.z, loz.z, hiz.z, chf.z, cnf.z, TC, NC)                      \
        CRT_WIDE_CHILD(lox.w, hix.w, loy.w, hiy.w, loz.w, hiz.w, chf.w, cnf.w, TD, ND)                      \
    }
// compare-exchange: swap when the second entry is strictly nearer
#define CRT_CX(TI, NI, TJ, NJ)              \
    {                                       \
        const bool sw = TJ < TI;            \
        const float ta = TI, tb = TJ;       \
        const int32_t na = NI, nb = NJ;     \
        TI = sw ? tb : ta;                  \
        TJ = sw ? ta : tb;                  \
        NI = sw ? nb : na;                  \
        NJ = sw ? na : nb;                  \
    }
        CRT_WIDE_GROUP(0, t0, n0, t1, n1, t2, n2, t3, n3)
        if constexpr (W == 8) CRT_WIDE_GROUP(1, t4, n4, t5, n5, t6, n6, t7, n7)
#undef CRT_WIDE_GROUP
#undef CRT_WIDE_CHILD
        // with leaves pending the first child is pushed too (on top)
        const int32_t pushes = np ? m : (m > 0 ? m - 1 : 0);
        if (sp + pushes > kCap) {
            status |= CRT_STATUS_STACK_OVERFLOW;
            np = 0;
            return kEnd;
        }
        if (any() && slot_order) {
            // slot order: the first hit child in slot order comes first, the
            // others are pushed last slot first
            int32_t f = -1, nf = -1;
            if constexpr (W == 8) {
                f = n7 >= 0 ? 7 : f;
                f = n6 >= 0 ? 6 : f;
                f = n5 >= 0 ? 5 : f;
                f = n4 >= 0 ? 4 : f;
            }
            f = n3 >= 0 ? 3 : f;
            f = n2 >= 0 ? 2 : f;
            f = n1 >= 0 ? 1 : f;
            f = n0 >= 0 ? 0 : f;
            const int32_t skip = np ? -1 : f;
            if constexpr (W == 8) {
                if (n7 >= 0 && skip != 7) st.put(sp++, n7, t7);
                if (n6 >= 0 && skip != 6) st.put(sp++, n6, t6);
                if (n5 >= 0 && skip != 5) st.put(sp++, n5, t5);
                if (n4 >= 0 && skip != 4) st.put(sp++, n4, t4);
                nf = f == 4 ? n4 : f == 5 ? n5 : f == 6 ? n6 : f == 7 ? n7 : nf;
            }
            if (n3 >= 0 && skip != 3) st.put(sp++, n3, t3);
            if (n2 >= 0 && skip != 2) st.put(sp++, n2, t2);
            if (n1 >= 0 && skip != 1) st.put(sp++, n1, t1);
            if (n0 >= 0 && skip != 0) st.put(sp++, n0, t0);
            nf = f == 0 ? n0 : f == 1 ? n1 : f == 2 ? n2 : f == 3 ? n3 : nf;
            if (np) {
                cur = -1;
                return kLeaves;
            }
            cur = nf;   // -1 (no inner child hit): pop first
            return kNext;
        }
        // nearest first: ascending entry distance by a fixed network, W = 4
        // five exchanges, W = 8 Batcher's odd-even merge sort (19 exchanges)
        if constexpr (W == 4) {
            CRT_CX(t0, n0, t1, n1) CRT_CX(t2, n2, t3, n3) CRT_CX(t0, n0, t2, n2) CRT_CX(t1, n1, t3, n3)
            CRT_CX(t1, n1, t2, n2)
        } else {
            CRT_CX(t0, n0, t1, n1) CRT_CX(t2, n2, t3, n3) CRT_CX(t4, n4, t5, n5) CRT_CX(t6, n6, t7, n7)
            CRT_CX(t0, n0, t2, n2) CRT_CX(t1, n1, t3, n3) CRT_CX(t4, n4, t6, n6) CRT_CX(t5, n5, t7, n7)
            CRT_CX(t1, n1, t2, n2) CRT_CX(t5, n5, t6, n6)
            CRT_CX(t0, n0, t4, n4) CRT_CX(t1, n1, t5, n5) CRT_CX(t2, n2, t6, n6) CRT_CX(t3, n3, t7, n7)
            CRT_CX(t2, n2, t4, n4) CRT_CX(t3, n3, t5, n5)
            CRT_CX(t1, n1, t2, n2) CRT_CX(t3, n3, t4, n4) CRT_CX(t5, n5, t6, n6)
        }
#undef CRT_CX
        // farthest first, so that pops come nearest first: sorted entry k (k < m) goes to stack
        // index sp + m - 1 - k; entry 0 only with leaves pending
        if (sp + m <= LDS_DEPTH) {
            // the whole run stays in LDS: entry k on row LDS_DEPTH - sp - m + k, all from one base
            LdsInt* b = st.row(LDS_DEPTH - sp - m);
            const int32_t rs = 2 * st.stride;   // one row
#define CRT_WIDE_PUT(K, N, T)                           \
    if (K < m) {                                        \
        b[K * rs] = N;                                  \
        b[K * rs + st.stride] = __float_as_int(T);      \
    }
            if constexpr (W == 8) {
                CRT_WIDE_PUT(7, n7, t7)
                CRT_WIDE_PUT(6, n6, t6)
                CRT_WIDE_PUT(5, n5, t5)
                CRT_WIDE_PUT(4, n4, t4)
            }
            CRT_WIDE_PUT(3, n3, t3)
            CRT_WIDE_PUT(2, n2, t2)
            CRT_WIDE_PUT(1, n1, t1)
            if (np) CRT_WIDE_PUT(0, n0, t0)
#undef CRT_WIDE_PUT
            sp += pushes;
            if (np) {
                cur = -1;
                return kLeaves;
            }
            cur = n0;   // -1 (no inner child hit): pop first
            return kNext;
        }
        const int32_t lowest = np ? 0 : 1;   // sorted index pushed last
        if (W == 8) {
            if (7 >= lowest && 7 < m) st.put(sp++, n7, t7);
            if (6 >= lowest && 6 < m) st.put(sp++, n6, t6);
            if (5 >= lowest && 5 < m) st.put(sp++, n5, t5);
            if (4 >= lowest && 4 < m) st.put(sp++, n4, t4);
        }
        if (3 >= lowest && 3 < m) st.put(sp++, n3, t3);
        if (2 >= lowest && 2 < m) st.put(sp++, n2, t2);
        if (1 < m) st.put(sp++, n1, t1);
        if (lowest == 0 && 0 < m) st.put(sp++, n0, t0);
        if (np) {
            cur = -1;
            return kLeaves;
        }
        cur = n0;   // -1 (no inner child hit): pop first
        return kNext;
    }

    // Leaf triangle `idx` (record loaded by the caller); true when an any-hit
    // walk is done.
    __device__ __forceinline__ bool test(int32_t idx, float4 a, float4 b, float4 c) {
        const float t = tri_t(ro, rd, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), mk(c.x, c.y, c.z));
        tests++;
        if (t > 0.0f && t < tmax) {
            tmax = t;
            tri = __float_as_int(a.w);
            mesh = __float_as_int(b.w);
            (void)idx;
            if (any()) return true;
        }
        return false;
    }

};

// One-ray-per-lane walk of one entity (the scene loop lives in thread_walk.h).
// Returns the node visits; hit in (t, tri, mesh) when one was accepted.
template <bool ANY, int W, int LDS_DEPTH>
__device__ __forceinline__ uint32_t walk_wide_thread(const SceneArgs& s, f3 ro, f3 rd, int32_t root, int32_t count,
                                                     float tmax, int32_t max_iter, float& out_t, int32_t& out_tri,
                                                     int32_t& out_mesh, uint32_t& tests, uint32_t& status,
                                                     int32_t* lds, int32_t stride, bool slot_order) {
    Core<ANY, W, LDS_DEPTH> c;
    Stack<LDS_DEPTH> st;
    int32_t spill_node[kSpill];
    float spill_t[kSpill];
    st.lds = lds;
    st.stride = stride;
    st.spill_node = spill_node;
    st.spill_t = spill_t;
    c.begin(ro, rd, tmax, root, count, s.total_nodes);
    c.tri = -1;
    c.mesh = -1;
    c.tests = 0;
    c.status = 0;
    const float4* __restrict__ rec = reinterpret_cast<const float4*>(s.tri_records);
    bool done = false;
    while (!done) {
        const int32_t r = c.step(s, max_iter, st, slot_order);
        if (r == kEnd) break;
        if (r == kLeaves) {
            for (int32_t j = 0; j < c.np && !done; ++j) {
                const int32_t leaf = st.leaf(j);
                const int32_t first = leaf >> 4, len = leaf & 0xF;
                for (int32_t idx = first; idx < first + len; ++idx) {
                    const float4* p = rec + 3 * int64_t(idx);
                    if (c.test(idx, p[0], p[1], p[2])) {
                        done = true;
                        break;
                    }
                }
            }
        }
    }
    out_t = c.tri >= 0 ? c.tmax : -1.0f;
    out_tri = c.tri;
    out_mesh = c.mesh;
    tests += c.tests;
    status |= c.status;
    return uint32_t(c.it);
}

// Wave-scheduled walk of up to 64 rays (one per lane) with the scene loop
// of IntersectScene (TraverseBVHStackless.glsl:280-366, :558-579):
//
//   node phase  lanes without pending leaves take node steps (up to INNER per
//               round); a lane whose step finds leaves parks;
//   leaf phase  once LEAF_BATCH lanes are parked (or no lane can step), every
//               parked lane tests its pending triangles, two records per
//               fetch, so the wave spends each triangle-test pass on as many
//               lanes as possible;
//   walk end    next entity (the walk starts from the scene loop's TMax, so a
//               hit the walk accepts always wins at scene level), or the hit
//               record.
template <bool ANY, int W, int LDS_DEPTH, int INNER, int LEAF_BATCH, bool DYN = false>
struct WideWave {
    Core<ANY, W, LDS_DEPTH, DYN> c;
    int32_t slot, ent, hit_ent;
    float ray_tmax;
    uint32_t visits;
    int32_t leaf, off;   // leaf being tested (start << 4 | count) and triangles done
    bool active, leafing, walk_done;

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
        c.begin(xform(en.inverse, mk(a.x, a.y, a.z), 1.0f), xform(en.inverse, mk(b.x, b.y, b.z), 0.0f),
                c.any() ? ray_tmax : c.tmax, en.node_offset, en.node_count, s.total_nodes);
        walk_done = false;
        return true;
    }

    __device__ __forceinline__ void finish(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                           crt_hit* __restrict__ hits, uint32_t* __restrict__ tri_tests,
                                           uint32_t flags) {
        float4 tuvw = make_float4(-1.0f, -1.0f, -1.0f, -1.0f);
        const int32_t tri = c.tri;
        const float closest = tri >= 0 ? c.tmax : -1.0f;
        if (c.any()) {
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
        o[1] = make_float4(__int_as_float(tri), __int_as_float(c.mesh), __int_as_float(tri >= 0 ? hit_ent : -1),
                           __uint_as_float(pack_status(visits, c.status)));
        if (tri_tests) tri_tests[slot] = c.tests;
    }

    // Lane takes ray i.
    __device__ __forceinline__ void start(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                          crt_hit* __restrict__ hits, int64_t i, uint32_t flags,
                                          uint32_t* __restrict__ tri_tests) {
        active = true;
        leafing = false;
        walk_done = false;
        slot = int32_t(i);
        ray_tmax = reinterpret_cast<const float4*>(rays + i)[0].w;
        ent = -1;
        hit_ent = -1;
        visits = 0;
        c.tmax = ray_tmax;
        c.tri = c.mesh = -1;
        c.tests = c.status = 0;
        if (!begin_walk(s, rays, flags)) {
            finish(s, rays, hits, tri_tests, flags);
            active = false;
        }
    }

    // One round of the wave: node phase, leaf phase, walk ends.
    __device__ __forceinline__ void round(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                          crt_hit* __restrict__ hits, uint32_t flags, int32_t max_iter,
                                          uint32_t* __restrict__ tri_tests, Stack<LDS_DEPTH>& st) {
        const float4* __restrict__ rec = reinterpret_cast<const float4*>(s.tri_records);
        const bool slot_order = (flags & CRT_FLAG_SLOT_ORDER) != 0;
        // --- node phase ----------------------------------------------------------
        for (int k = 0; k < INNER; ++k) {
            if (active && !walk_done && !leafing) {
                const int32_t r = c.step(s, max_iter, st, slot_order);
                if (r == kEnd) {
                    walk_done = true;
                } else if (r == kLeaves) {
                    leafing = true;
                    leaf = st.leaf(0);
                    c.pi = 1;
                    off = 0;
                }
            }
            const unsigned long long stepping = __ballot(active && !walk_done && !leafing);
            if (stepping == 0 || __popcll(__ballot(leafing)) >= LEAF_BATCH) break;
        }
        // --- leaf phase: pending triangles, two records per fetch ----------------
        // Straight-line code for every lane of the wave, the state updated by
        // selects: a lane that is not testing reads record 0 (which exists
        // whenever some lane has a leaf) and a clamped leaf slot, and its
        // results are discarded.  Branches around the tests would make the
        // compiler copy the loop-carried state on every iteration.
        while (__ballot(leafing)) {
            const bool act = leafing;
            const int32_t len = leaf & 0xF;
            const int32_t first = (leaf >> 4) + off, left = len - off;
            const bool two = act && left > 1;   // an idle lane reads record 0 only
            const float4* p = rec + 3 * int64_t(act ? first : 0);
            // the second record, or the first again when one is left (never past the leaf)
            const float4* p1 = p + (two ? 3 : 0);
            const float4 a0 = p[0], b0 = p[1], c0 = p[2];
            const float4 a1 = p1[0], b1 = p1[1], c1 = p1[2];
            const float ta = tri_t(c.ro, c.rd, mk(a0.x, a0.y, a0.z), mk(b0.x, b0.y, b0.z), mk(c0.x, c0.y, c0.z));
            const float tb = tri_t(c.ro, c.rd, mk(a1.x, a1.y, a1.z), mk(b1.x, b1.y, b1.z), mk(c1.x, c1.y, c1.z));
            // first record, then the second unless the first ended an any-hit walk
            // (acceptance 0 < t < tmax, the first found wins a tie: Core::test)
            const bool acc_a = act && ta > 0.0f && ta < c.tmax;
            const float tm = acc_a ? ta : c.tmax;
            const bool test_b = act && two && !(c.any() && acc_a);
            const bool acc_b = test_b && tb > 0.0f && tb < tm;
            c.tmax = acc_b ? tb : tm;
            c.tri = acc_b ? __float_as_int(a1.w) : (acc_a ? __float_as_int(a0.w) : c.tri);
            c.mesh = acc_b ? __float_as_int(b1.w) : (acc_a ? __float_as_int(b0.w) : c.mesh);
            hit_ent = (acc_a || acc_b) ? ent : hit_ent;
            c.tests += act ? (test_b ? 2u : 1u) : 0u;
            const bool done = c.any() && (acc_a || acc_b);
            // next: more of this leaf, the next pending leaf, or the end of the phase
            const int32_t off2 = off + (two ? 2 : 1);
            const bool leaf_end = off2 >= len;
            const bool more = c.pi < c.np;
            const int32_t nl = st.leaf(uint32_t(c.pi) < uint32_t(W) ? c.pi : W - 1);
            const bool adv = act && !done && leaf_end && more;
            leaf = adv ? nl : leaf;
            c.pi += adv ? 1 : 0;
            off = adv ? 0 : (act ? off2 : off);
            walk_done = walk_done || (act && done);
            leafing = act && !done && !(leaf_end && !more);
        }
        // --- walk ends: next entity or the ray's result --------------------------
        // the ends run once kWalkEndBatch lanes wait or no lane can step
        const unsigned long long ends = __ballot(walk_done);
        if (ends && (__popcll(ends) >= kWalkEndBatch || !__ballot(active && !walk_done && !leafing))) {
            if (walk_done) {
                walk_done = false;
                visits += uint32_t(c.it);
                bool fin = c.any() && c.tri >= 0;
                if (!fin) fin = !begin_walk(s, rays, flags);
                if (fin) {
                    finish(s, rays, hits, tri_tests, flags);
                    active = false;
                }
            }
        }
    }

    // Rays i of 64 lanes (i < 0: lane idle) until all are done.
    __device__ __forceinline__ void run(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                        crt_hit* __restrict__ hits, int64_t i, uint32_t flags, int32_t max_iter,
                                        uint32_t* __restrict__ tri_tests, Stack<LDS_DEPTH> st) {
        active = false;
        leafing = false;
        walk_done = false;
        if (i >= 0) start(s, rays, hits, i, flags, tri_tests);
        while (__ballot(active)) round(s, rays, hits, flags, max_iter, tri_tests, st);
    }

    // Rays [0, n) claimed from the sharded queue (queue.h) by the whole wave:
    // between rounds, once at least `refill` lanes are idle, one claim takes a
    // ray for each idle lane (lane ranks by mbcnt), so a wave keeps its lanes
    // busy instead of waiting for its longest ray.  Ends when the queue is dry
    // and every lane is done.
    __device__ __forceinline__ void run_refill(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                               crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                               int32_t max_iter, uint32_t* __restrict__ tri_tests,
                                               Stack<LDS_DEPTH> st, unsigned long long* __restrict__ counters,
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
                        if (rank < got) start(s, rays, hits, q.at(first + rank), flags, tri_tests);
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

}  // namespace wide
}  // namespace crt
