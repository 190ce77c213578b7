// Speculative wave walk of the wide layout (BVH4 / BVH8): lanes never wait
// for their leaf tests, and the wave tests leaves 64 triangles at a time.
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
//              the hit leaves' triangles go to the wave's item list;
//   items      each item is one triangle test, (record << 6 | owner lane),
//              appended in lane order, then slot order, then triangle order
//              (an exclusive scan of the per-lane counts by ballot bit planes);
//              a lane whose triangles do not fit in the list waits (parked)
//              until the next leaf phase;
//   leaf phase after INNER node steps, once LIST_TRIGGER items wait or when no
//              lane can step: item i of the list is tested by lane i % 64
//              with the owner's ray (lane shuffles); an accepted test
//              (0 < t < the owner's tmax) enters a 64-bit LDS atomic minimum
//              per owner: (t, item) for closest hits, (item, t) for any hits;
//              then each owner takes its winner (closest: nearer hit; any:
//              the walk ends);
//   culling    box tests and pops use the tmax of the last leaf phase, which
//              only errs toward visiting more nodes, so a closest-hit walk
//              still reports the nearest triangle that any wide walk finds.
//
// A lane's visits and tests depend on the leaf phases, which depend on the
// rays sharing its wave: with the grid launch (64 consecutive rays per wave)
// the records are deterministic and the host twin (oracle_spec.c) replays the
// schedule; with the refilling launch the nearest t, the hit / miss answer
// and any-hit occlusion are still fixed, the triangle is fixed up to exact
// distance ties, and the counters vary.
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

// LDS of one wave: its item list and the 64 owners' keys.
template <int LIST>
struct WaveLds {
    uint32_t items[LIST];
    unsigned long long keys[64];
};

__device__ __forceinline__ int32_t lanes_below(unsigned long long mask) {
    return int32_t(__builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u)));
}

// An item packs a triangle record index above the 6 owner-lane bits.
constexpr int32_t kMaxRecord = 1 << 26;

template <bool ANY, int W, int LDS_DEPTH, int INNER, int LIST_TRIGGER, int LIST>
struct SpecWave {
    static constexpr int G = W / 4;
    static constexpr int32_t kCap = LDS_DEPTH + wide::kSpill;
    static_assert(LIST >= LIST_TRIGGER && LIST >= W * 15, "a lane's leaves of one step always fit an empty list");

    // per lane
    f3 ro, rd, inv;
    float tmax, ray_tmax;
    int32_t root, count, cur, sp, it;
    int32_t tri, mesh, slot, ent, hit_ent;
    uint32_t tests, status, visits;
    int32_t np;   // leaves of the last step waiting for the list (in the lane's LDS column)
    bool active, walk_done, parked, inlist;
    // per wave (uniform)
    int32_t nitems;

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
        if (ANY) tmax = ray_tmax;
        root = en.node_offset;
        count = en.node_count;
        cur = root;
        sp = 0;
        it = 0;
        np = 0;
        walk_done = false;
        parked = false;
        inlist = false;
        return true;
    }

    __device__ __forceinline__ void finish(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                           crt_hit* __restrict__ hits, uint32_t* __restrict__ tri_tests,
                                           uint32_t flags) {
        float4 tuvw = make_float4(-1.0f, -1.0f, -1.0f, -1.0f);
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

    // Pop (culling by the current tmax), guards and one node; the node's hit
    // leaves go to the lane's LDS column (np of them).  False when the walk
    // is over.
    __device__ __forceinline__ bool step(const SceneArgs& s, int32_t max_iter, Stack<LDS_DEPTH>& st) {
        if (cur < 0) {
            bool found = false;
            while (sp > 0) {
                --sp;
                int32_t node;
                float t;
                st.get(sp, node, t);
                if (!ANY && t > tmax * kWideGrow) continue;
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
        if (cur < root || cur >= root + count || int64_t(cur) >= s.total_nodes) {
            status |= CRT_STATUS_BAD_INDEX;
            return false;
        }
        it++;
        int32_t nl = 0;
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
        if (hit && cnt > 0) {                                                                               \
            if (chi >= 0 && chi + cnt <= kMaxRecord) {                                                      \
                st.put_leaf(nl, (chi << 4) | cnt);                                                          \
                ++nl;                                                                                       \
            } else {                                                                                        \
                status |= CRT_STATUS_BAD_INDEX;   /* not representable in an item */                        \
            }                                                                                               \
        }                                                                                                   \
        const bool inner = hit && cnt == 0;                                                                 \
        T = inner ? e0 : inf;                                                                               \
        N = inner ? chi : -1;                                                                               \
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
        const int32_t m = (n0 >= 0) + (n1 >= 0) + (n2 >= 0) + (n3 >= 0) + (n4 >= 0) + (n5 >= 0) + (n6 >= 0) +
                          (n7 >= 0);
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
        // farthest first, so that pops come nearest first; the nearest is next
        if constexpr (W == 8) {
            if (7 < m) st.put(sp++, n7, t7);
            if (6 < m) st.put(sp++, n6, t6);
            if (5 < m) st.put(sp++, n5, t5);
            if (4 < m) st.put(sp++, n4, t4);
        }
        if (3 < m) st.put(sp++, n3, t3);
        if (2 < m) st.put(sp++, n2, t2);
        if (1 < m) st.put(sp++, n1, t1);
        cur = n0;   // -1 (no inner child hit): pop first
        np = nl;
        return true;
    }

    // Appends the waiting leaves of every lane, in lane order, as far as the
    // list has room; a lane that does not fit stays parked.
    __device__ __forceinline__ void append(Stack<LDS_DEPTH>& st, WaveLds<LIST>& L) {
        int32_t n = 0;
        for (int32_t j = 0; j < np; ++j) n += st.leaf(j) & 0xF;
        // exclusive scan of n over the wave by bit planes
        int32_t excl = 0;
        for (int b = 0; b < 8; ++b) {
            const unsigned long long plane = __ballot((n >> b) & 1);
            if (plane == 0ull && __ballot(n >> b) == 0ull) break;
            excl += lanes_below(plane) << b;
        }
        const int32_t room = LIST - nitems;
        const bool fits = n > 0 && excl + n <= room;
        const unsigned long long fit_mask = __ballot(fits);
        if (fits) {
            int32_t k = nitems + excl;
            const uint32_t lane = uint32_t(__lane_id());
            for (int32_t j = 0; j < np; ++j) {
                const int32_t leaf = st.leaf(j);
                const int32_t first = leaf >> 4, len = leaf & 0xF;
                for (int32_t x = 0; x < len; ++x) L.items[k++] = (uint32_t(first + x) << 6) | lane;
            }
            tests += uint32_t(n);
            np = 0;
            parked = false;
            inlist = true;
        } else if (n > 0) {
            parked = true;
        }
        if (fit_mask) {
            const int32_t last = 63 - __builtin_clzll(fit_mask);
            nitems += __shfl(excl + n, last);
        }
    }

    // Tests every item of the list, 64 at a time, and gives each owner its winner.
    __device__ __forceinline__ void leaf_phase(const SceneArgs& s, WaveLds<LIST>& L) {
        const int32_t lane = __lane_id();
        L.keys[lane] = ~0ull;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        const float4* __restrict__ rec = reinterpret_cast<const float4*>(s.tri_records);
        for (int32_t base = 0; base < nitems; base += 64) {
            const int32_t idx = base + lane;
            const bool valid = idx < nitems;
            const uint32_t item = valid ? L.items[idx] : 0u;
            const int32_t owner = int32_t(item & 63u);
            const float ox = __shfl(ro.x, owner), oy = __shfl(ro.y, owner), oz = __shfl(ro.z, owner);
            const float dx = __shfl(rd.x, owner), dy = __shfl(rd.y, owner), dz = __shfl(rd.z, owner);
            const float otmax = __shfl(tmax, owner);
            if (valid) {
                const float4* p = rec + 3 * int64_t(item >> 6);
                const float4 a = p[0], b = p[1], c = p[2];
                const float t = tri_t(mk(ox, oy, oz), mk(dx, dy, dz), mk(a.x, a.y, a.z), mk(b.x, b.y, b.z),
                                      mk(c.x, c.y, c.z));
                if (t > 0.0f && t < otmax) {
                    const unsigned long long tb = __float_as_uint(t), ib = uint32_t(idx);
                    atomicMin(&L.keys[owner], ANY ? ((ib << 32) | tb) : ((tb << 32) | ib));
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const unsigned long long k = L.keys[lane];
        if (k != ~0ull && inlist) {
            const uint32_t ib = ANY ? uint32_t(k >> 32) : uint32_t(k);
            const uint32_t tb = ANY ? uint32_t(k) : uint32_t(k >> 32);
            const float4* p = rec + 3 * int64_t(L.items[ib] >> 6);
            tmax = __uint_as_float(tb);
            tri = __float_as_int(p[0].w);
            mesh = __float_as_int(p[1].w);
            hit_ent = ent;
            if (ANY) {   // the first accepted triangle ends an any-hit walk
                walk_done = true;
                np = 0;
                parked = false;
            }
        }
        inlist = false;
        nitems = 0;
    }

    // One round of the wave: node steps, leaf phase, walk ends.
    __device__ __forceinline__ void round(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                          crt_hit* __restrict__ hits, uint32_t flags, int32_t max_iter,
                                          uint32_t* __restrict__ tri_tests, Stack<LDS_DEPTH>& st,
                                          WaveLds<LIST>& L) {
        for (int k = 0; k < INNER; ++k) {
            if (active && !walk_done && !parked) {
                if (!step(s, max_iter, st)) walk_done = true;
            }
            append(st, L);
            const unsigned long long stepping = __ballot(active && !walk_done && !parked);
            if (stepping == 0 || nitems >= LIST_TRIGGER) break;
        }
        if (nitems > 0) {
            leaf_phase(s, L);
            append(st, L);   // parked lanes' leaves go first into the empty list
        }
        // a walk ends once it has no step left and no triangle waiting
        const bool ends = active && walk_done && !parked && !inlist;
        if (__ballot(ends)) {
            if (ends) {
                visits += uint32_t(it);
                bool fin = ANY && tri >= 0;
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
                                        uint32_t* __restrict__ tri_tests, Stack<LDS_DEPTH> st, WaveLds<LIST>& L) {
        active = false;
        walk_done = parked = inlist = false;
        np = 0;
        nitems = 0;
        if (i >= 0) begin_ray(s, rays, hits, i, flags, tri_tests);
        while (__ballot(active)) round(s, rays, hits, flags, max_iter, tri_tests, st, L);
    }

    // Rays [0, n) claimed from the sharded queue between rounds, one per idle
    // lane, once at least `refill` lanes are idle.
    __device__ __forceinline__ void run_refill(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                               crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                               int32_t max_iter, uint32_t* __restrict__ tri_tests,
                                               Stack<LDS_DEPTH> st, WaveLds<LIST>& L,
                                               unsigned long long* __restrict__ counters, int32_t refill) {
        active = false;
        walk_done = parked = inlist = false;
        np = 0;
        nitems = 0;
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
