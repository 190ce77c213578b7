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
//             the exit distance and tmax grown by kWideGrow; in slot order a
//             hit leaf is appended to the pending list and a hit inner child
//             competes for "near" (strictly smaller entry distance wins; the
//             loser is pushed)
//   leaves    with pending leaves the near child is pushed too, the pending
//             triangles are tested in slot order, and the next node is popped
//             (with the tmax those tests produced); without, the walk goes to
//             the near child, else pops
//   overflow  a push beyond LDS_DEPTH + 64 entries ends the walk with
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

    __device__ __forceinline__ void put_leaf(int32_t j, int32_t v) { lds[(2 * LDS_DEPTH + j) * stride] = v; }
    __device__ __forceinline__ int32_t leaf(int32_t j) const { return lds[(2 * LDS_DEPTH + j) * stride]; }

    __device__ __forceinline__ void put(int32_t sp, int32_t node, float t) {
        if (sp < LDS_DEPTH) {
            lds[(2 * sp) * stride] = node;
            lds[(2 * sp + 1) * stride] = __float_as_int(t);
        } else {
            spill_node[sp - LDS_DEPTH] = node;
            spill_t[sp - LDS_DEPTH] = t;
        }
    }
    __device__ __forceinline__ void get(int32_t sp, int32_t& node, float& t) const {
        if (sp < LDS_DEPTH) {
            node = lds[(2 * sp) * stride];
            t = __int_as_float(lds[(2 * sp + 1) * stride]);
        } else {
            node = spill_node[sp - LDS_DEPTH];
            t = spill_t[sp - LDS_DEPTH];
        }
    }
};

// One ray's walk of one entity's wide tree.
template <bool ANY, int W, int LDS_DEPTH>
struct Core {
    static constexpr int G = W / 4;
    static constexpr int32_t kCap = LDS_DEPTH + kSpill;

    f3 ro, rd, inv;
    float tmax;          // closest: shrinks with every accepted hit
    int32_t root, count;
    int32_t cur;         // next node, or -1: pop first
    int32_t sp, it;
    int32_t np, pi;      // pending leaves (in LDS, start << 4 | count): count, next to test
    int32_t tri, mesh;   // accepted hit of this walk
    uint32_t tests, status;

    __device__ __forceinline__ void begin(f3 lro, f3 lrd, float walk_tmax, int32_t r, int32_t c) {
        ro = lro;
        rd = lrd;
        inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
        tmax = walk_tmax;
        root = r;
        count = c;
        cur = r;
        sp = 0;
        it = 0;
        np = 0;
        pi = 0;
    }

    // Pop, guards, one node.  kNext: cur holds the next node; kLeaves: np
    // pending leaves to test before the next step; kEnd: the walk is over.
    __device__ __forceinline__ int32_t step(const SceneArgs& s, int32_t max_iter, Stack<LDS_DEPTH>& st) {
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
            if (!found) return kEnd;
        }
        if (it >= max_iter) {
            status |= CRT_STATUS_ITER_CAP;
            return kEnd;
        }
        if (cur < root || cur >= root + count || int64_t(cur) >= s.total_nodes) {
            status |= CRT_STATUS_BAD_INDEX;
            return kEnd;
        }
        it++;
        np = 0;
        pi = 0;
        const float4* __restrict__ groups = reinterpret_cast<const float4*>(s.nodes);
        int32_t near_node = -1;
        float near_t = 0.0f;
        bool overflow = false;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const float4* p = groups + (int64_t(cur) * G + g) * 8;
            const float4 lox = p[0], hix = p[1], loy = p[2], hiy = p[3], loz = p[4], hiz = p[5];
            const float4 chf = p[6], cnf = p[7];
            const float lx[4] = {lox.x, lox.y, lox.z, lox.w}, hx[4] = {hix.x, hix.y, hix.z, hix.w};
            const float ly[4] = {loy.x, loy.y, loy.z, loy.w}, hy[4] = {hiy.x, hiy.y, hiy.z, hiy.w};
            const float lz[4] = {loz.x, loz.y, loz.z, loz.w}, hz[4] = {hiz.x, hiz.y, hiz.z, hiz.w};
            const int32_t ch[4] = {__float_as_int(chf.x), __float_as_int(chf.y), __float_as_int(chf.z),
                                   __float_as_int(chf.w)};
            const int32_t cn[4] = {__float_as_int(cnf.x), __float_as_int(cnf.y), __float_as_int(cnf.z),
                                   __float_as_int(cnf.w)};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float fx = (hx[k] - ro.x) * inv.x, fy = (hy[k] - ro.y) * inv.y, fz = (hz[k] - ro.z) * inv.z;
                const float nx = (lx[k] - ro.x) * inv.x, ny = (ly[k] - ro.y) * inv.y, nz = (lz[k] - ro.z) * inv.z;
                const float t1 = mn(mn(mx(fx, nx), mn(mx(fy, ny), mx(fz, nz))) * kWideGrow, tmax * kWideGrow);
                const float t0 = mx(mx(mn(fx, nx), mx(mn(fy, ny), mn(fz, nz))), 0.0f);
                const bool hit = cn[k] >= 0 && t1 >= t0 && (t0 > 0.0f || t1 > 0.0f);
                if (hit && cn[k] > 0) {
                    st.put_leaf(np, (ch[k] << 4) | cn[k]);
                    ++np;
                } else if (hit) {
                    int32_t push_node = ch[k];
                    float push_t = t0;
                    bool push = true;
                    if (near_node < 0) {
                        near_node = ch[k];
                        near_t = t0;
                        push = false;
                    } else if (t0 < near_t) {
                        push_node = near_node;
                        push_t = near_t;
                        near_node = ch[k];
                        near_t = t0;
                    }
                    if (push) {
                        if (sp >= kCap) {
                            overflow = true;
                        } else {
                            st.put(sp, push_node, push_t);
                            ++sp;
                        }
                    }
                }
            }
        }
        if (overflow) {
            status |= CRT_STATUS_STACK_OVERFLOW;
            np = 0;
            return kEnd;
        }
        if (np) {
            if (near_node >= 0) {
                if (sp >= kCap) {
                    status |= CRT_STATUS_STACK_OVERFLOW;
                    np = 0;
                    return kEnd;
                }
                st.put(sp, near_node, near_t);
                ++sp;
            }
            cur = -1;
            return kLeaves;
        }
        cur = near_node;   // -1 (none hit): pop first
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
            if (ANY) return true;
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
                                                     int32_t* lds, int32_t stride) {
    Core<ANY, W, LDS_DEPTH> c;
    Stack<LDS_DEPTH> st;
    int32_t spill_node[kSpill];
    float spill_t[kSpill];
    st.lds = lds;
    st.stride = stride;
    st.spill_node = spill_node;
    st.spill_t = spill_t;
    c.begin(ro, rd, tmax, root, count);
    c.tri = -1;
    c.mesh = -1;
    c.tests = 0;
    c.status = 0;
    const float4* __restrict__ rec = reinterpret_cast<const float4*>(s.tri_records);
    bool done = false;
    while (!done) {
        const int32_t r = c.step(s, max_iter, st);
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
template <bool ANY, int W, int LDS_DEPTH, int INNER, int LEAF_BATCH>
struct WideWave {
    Core<ANY, W, LDS_DEPTH> c;
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
                ANY ? ray_tmax : c.tmax, en.node_offset, en.node_count);
        walk_done = false;
        return true;
    }

    __device__ __forceinline__ void finish(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                           crt_hit* __restrict__ hits, uint32_t* __restrict__ tri_tests,
                                           uint32_t flags) {
        float4 tuvw = make_float4(-1.0f, -1.0f, -1.0f, -1.0f);
        const int32_t tri = c.tri;
        const float closest = tri >= 0 ? c.tmax : -1.0f;
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
        o[1] = make_float4(__int_as_float(tri), __int_as_float(c.mesh), __int_as_float(tri >= 0 ? hit_ent : -1),
                           __uint_as_float(pack_status(visits, c.status)));
        if (tri_tests) tri_tests[slot] = c.tests;
    }

    __device__ __forceinline__ void run(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                        crt_hit* __restrict__ hits, int64_t i, uint32_t flags, int32_t max_iter,
                                        uint32_t* __restrict__ tri_tests, Stack<LDS_DEPTH> st) {
        const float4* __restrict__ rec = reinterpret_cast<const float4*>(s.tri_records);
        active = i >= 0;
        leafing = false;
        walk_done = false;
        if (active) {
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
        while (__ballot(active)) {
            // --- node phase ------------------------------------------------------
            for (int k = 0; k < INNER; ++k) {
                if (active && !walk_done && !leafing) {
                    const int32_t r = c.step(s, max_iter, st);
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
            // --- leaf phase: pending triangles, two records per fetch ------------
            while (__ballot(leafing)) {
                if (leafing) {
                    const int32_t first = (leaf >> 4) + off, left = (leaf & 0xF) - off;
                    const float4* p = rec + 3 * int64_t(first);
                    const float4 a0 = p[0], b0 = p[1], c0 = p[2];
                    float4 a1 = a0, b1 = b0, c1 = c0;
                    if (left > 1) {
                        a1 = p[3];
                        b1 = p[4];
                        c1 = p[5];
                    }
                    const float before = c.tmax;
                    bool done = c.test(first, a0, b0, c0);
                    if (!done && left > 1) done = c.test(first + 1, a1, b1, c1);
                    if (c.tmax < before) hit_ent = ent;
                    off += left > 1 ? 2 : 1;
                    if (done) {
                        leafing = false;
                        walk_done = true;
                    } else if (off >= (leaf & 0xF)) {
                        if (c.pi < c.np) {
                            leaf = st.leaf(c.pi);
                            ++c.pi;
                            off = 0;
                        } else {
                            leafing = false;
                        }
                    }
                }
            }
            // --- walk ends: next entity or the ray's result ----------------------
            if (__ballot(walk_done)) {
                if (walk_done) {
                    walk_done = false;
                    visits += uint32_t(c.it);
                    bool fin = ANY && c.tri >= 0;
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

}  // namespace wide
}  // namespace crt
