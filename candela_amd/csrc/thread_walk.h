// One ray per lane: the reference's walks and scene loop as plain device
// functions (used by the one-ray-per-lane kernels of trace.hip and inline by
// the shading kernels of shade.hip).  Each function restates one GLSL
// function of Source/Core/Shaders/Intersectors/Include/:
//   walk_stackless<ANY>  TraverseBVHStackless.glsl:175-278 / :463-556
//   walk_stack<ANY>      TraverseBVHStack.glsl:168-324 / :509-657
//   wide::walk_wide_thread  the same closest / any-hit contract over the wide
//                        layout derived from either (wide_walk.h)
//   trace_one            IntersectScene / IntersectSceneIgnoreTransparent
//                        (TraverseBVHStackless.glsl:280-366) and the
//                        occlusion IntersectScene (:558-579)
// The stack walk keeps its first LDS_DEPTH entries in LDS (entry-major, so
// lane i of a wave touches bank i) and the rest in a per-lane scratch array.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"
#include "device_math.h"
#include "trace_launch.h"
#include "wide_walk.h"

namespace crt {
namespace dev {

using namespace ::crt::dm;

struct Walk {
    float t;
    int32_t tri;
    int32_t mesh;
    uint32_t tests;  // triangle tests (summed over the scene loop)
};

// Tests triangles [start, start+len).  Returns true when an any-hit walk is done.
template <bool ANY>
__device__ __forceinline__ bool test_leaf(const SceneArgs& s, f3 ro, f3 rd, int32_t start, int32_t len, float& tmax,
                                          Walk& w) {
    for (int32_t idx = start; idx < start + len; ++idx) {
        const int4 tr = *reinterpret_cast<const int4*>(s.tris[idx].packed);
        const f3 a = vertex_pos(s.verts, tr.x), b = vertex_pos(s.verts, tr.y), c = vertex_pos(s.verts, tr.z);
        const float t = tri_t(ro, rd, a, b, c);
        w.tests++;
        if (t > 0.0f && t < tmax) {
            tmax = t;
            w.t = t;
            w.mesh = tr.w;
            w.tri = idx;
            if (ANY) return true;
        }
    }
    return false;
}

// IntersectBVHStackless / IntersectBVHStacklessOcclusion.  Returns the node
// visits; `status` collects CRT_STATUS_* bits.
template <bool ANY>
__device__ __forceinline__ uint32_t walk_stackless(const SceneArgs& s, f3 ro, f3 rd, int32_t start, int32_t count,
                                                   float tmax, int32_t max_iter, Walk& w, uint32_t& status) {
    const float4* __restrict__ nodes = reinterpret_cast<const float4*>(s.nodes);
    const f3 inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    int32_t it = 0;
    int32_t ptr = start;
    bool stopped = false;
    w.t = -1.0f;
    w.tri = -1;
    w.mesh = -1;
    while (ptr >= 0 && it < max_iter) {
        if (ptr < start || ptr > start + count || int64_t(ptr) >= s.total_nodes) {
            status |= CRT_STATUS_BAD_INDEX;
            stopped = true;
            break;
        }
        it++;
        const float4 nmin = nodes[2 * int64_t(ptr)];
        const float4 nmax = nodes[2 * int64_t(ptr) + 1];
        const float tb = bounds_stackless(nmin, nmax, ro, inv, tmax);
        const int32_t link = __float_as_int(nmax.w);
        if (tb > 0.0f && tb < tmax) {
            const int32_t packed = __float_as_int(nmin.w);
            if (packed != -1) {
                if (test_leaf<ANY>(s, ro, rd, packed >> 4, packed & 0xF, tmax, w)) return uint32_t(it);
            } else {
                ptr++;
                continue;
            }
        }
        if (link < 0) {
            stopped = true;
            break;
        }
        ptr = link + start;
    }
    if (!stopped && ptr >= 0) status |= CRT_STATUS_ITER_CAP;
    return uint32_t(it);
}

// IntersectBVHStack / IntersectBVHStackOcclusion with the first LDS_DEPTH
// stack entries in LDS (`lds` points at this lane's column, stride
// `lds_stride` ints) and the rest in per-lane scratch.
template <bool ANY, int LDS_DEPTH>
__device__ __forceinline__ uint32_t walk_stack(const SceneArgs& s, f3 ro, f3 rd, int32_t start, int32_t count,
                                               float tmax, int32_t max_iter, Walk& w, uint32_t& status, int32_t* lds,
                                               int32_t lds_stride) {
    const float4* __restrict__ nodes = reinterpret_cast<const float4*>(s.nodes);
    const f3 inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    constexpr int kSpill = (64 - LDS_DEPTH) > 0 ? (64 - LDS_DEPTH) : 1;
    int32_t spill[kSpill];
    LdsInt* const col = (LdsInt*)lds;   // see LdsInt
    int32_t sp = 0;
    int32_t it = 0;
    int32_t cur = start;
    bool stopped = false;
    w.t = -1.0f;
    w.tri = -1;
    w.mesh = -1;
    while (it < max_iter) {
        if (cur < start || cur > start + count || cur < 0 || int64_t(cur) >= s.total_nodes) {
            status |= CRT_STATUS_BAD_INDEX;
            stopped = true;
            break;
        }
        it++;
        const float4* np = nodes + 4 * int64_t(cur);
        const float4 lmin = np[0], lmax = np[1], rmin = np[2], rmax = np[3];
        const int32_t lp = __float_as_int(lmin.w), rp = __float_as_int(rmin.w);
        const bool lleaf = lp != -1, rleaf = rp != -1;
        const float lt = lleaf ? -1.0f : bounds_stack(lmin, lmax, ro, inv, tmax);
        const float rt = rleaf ? -1.0f : bounds_stack(rmin, rmax, ro, inv, tmax);
        if (lleaf && test_leaf<ANY>(s, ro, rd, lp >> 4, lp & 0xF, tmax, w)) return uint32_t(it);
        if (rleaf && test_leaf<ANY>(s, ro, rd, rp >> 4, rp & 0xF, tmax, w)) return uint32_t(it);
        if (lt > 0.0f && rt > 0.0f) {
            int32_t next = __float_as_int(lmax.w) + start;
            int32_t post = __float_as_int(rmax.w) + start;
            if (rt < lt) {
                const int32_t tmp = next;
                next = post;
                post = tmp;
            }
            cur = next;
            if (sp >= 63) {
                status |= CRT_STATUS_STACK_OVERFLOW;
                stopped = true;
                break;
            }
            if (sp < LDS_DEPTH)
                col[sp * lds_stride] = post;
            else
                spill[sp - LDS_DEPTH] = post;
            sp++;
            continue;
        } else if (lt > 0.0f) {
            cur = __float_as_int(lmax.w) + start;
            continue;
        } else if (rt > 0.0f) {
            cur = __float_as_int(rmax.w) + start;
            continue;
        }
        if (sp <= 0) {
            stopped = true;
            break;
        }
        --sp;
        cur = sp < LDS_DEPTH ? col[sp * lds_stride] : spill[sp - LDS_DEPTH];
    }
    if (!stopped) status |= CRT_STATUS_ITER_CAP;
    return uint32_t(it);
}

template <int FORMAT, bool ANY, int LDS_DEPTH>
__device__ __forceinline__ uint32_t walk_entity(const SceneArgs& s, f3 ro, f3 rd, int32_t start, int32_t count,
                                                float tmax, int32_t max_iter, Walk& w, uint32_t& status,
                                                int32_t* lds, int32_t lds_stride, bool slot_order) {
    if constexpr (FORMAT == CRT_FORMAT_STACKLESS)
        return walk_stackless<ANY>(s, ro, rd, start, count, tmax, max_iter, w, status);
    else if constexpr (FORMAT == CRT_FORMAT_STACK)
        return walk_stack<ANY, LDS_DEPTH>(s, ro, rd, start, count, tmax, max_iter, w, status, lds, lds_stride);
    else
        return wide::walk_wide_thread<ANY, FORMAT == CRT_FORMAT_WIDE8 ? 8 : 4, LDS_DEPTH>(
            s, ro, rd, start, count, tmax, max_iter, w.t, w.tri, w.mesh, w.tests, status, lds, lds_stride,
            slot_order);
}


// The per-ray scene loop: IntersectScene / IntersectSceneIgnoreTransparent
// (closest) or the occlusion IntersectScene (any).  Returns the hit record;
// `tests` receives the triangle tests.
template <int FORMAT, bool ANY, int LDS_DEPTH>
__device__ __forceinline__ crt_hit scene_trace(const SceneArgs& s, const crt_ray& ray, uint32_t flags,
                                               int32_t max_iter, uint32_t& tests, int32_t* lds,
                                               int32_t lds_stride) {
    const f3 ro = mk(ray.org[0], ray.org[1], ray.org[2]);
    const f3 rd = mk(ray.dir[0], ray.dir[1], ray.dir[2]);
    uint32_t visits = 0, status = 0;
    float4 tuvw = make_float4(-1.0f, -1.0f, -1.0f, -1.0f);
    int32_t tri = -1, mesh = -1, ent = -1;
    Walk w;
    w.tests = 0;
    if constexpr (ANY) {
        for (int32_t i = 0; i < s.entity_count; ++i) {
            const crt_entity& e = s.entities[i];
            if ((flags & CRT_FLAG_IGNORE_TRANSPARENT) && __int_as_float(e.data[1]) < 0.99f) continue;
            const f3 lro = xform(e.inverse, ro, 1.0f);
            const f3 lrd = xform(e.inverse, rd, 0.0f);
            visits += walk_entity<FORMAT, true, LDS_DEPTH>(s, lro, lrd, e.node_offset, e.node_count, ray.tmax,
                                                           max_iter, w, status, lds, lds_stride,
                                                           (flags & CRT_FLAG_SLOT_ORDER) != 0);
            if (w.t > 0.0f) {
                tuvw.x = w.t;
                tri = w.tri;
                mesh = w.mesh;
                ent = i;
                break;
            }
        }
    } else {
        float closest = -1.0f;
        float tmax = ray.tmax;
        for (int32_t i = 0; i < s.entity_count; ++i) {
            const crt_entity& e = s.entities[i];
            if ((flags & CRT_FLAG_IGNORE_TRANSPARENT) && __int_as_float(e.data[1]) < 0.99f) continue;
            const f3 lro = xform(e.inverse, ro, 1.0f);
            const f3 lrd = xform(e.inverse, rd, 0.0f);
            visits += walk_entity<FORMAT, false, LDS_DEPTH>(s, lro, lrd, e.node_offset, e.node_count, tmax,
                                                            max_iter, w, status, lds, lds_stride, false);
            if (w.t > 0.0f && w.t < tmax) {
                tmax = w.t;
                closest = w.t;
                mesh = w.mesh;
                tri = w.tri;
                ent = i;
            }
        }
        const bool ok = closest > 0.0f && ((flags & CRT_FLAG_TRI0_IS_MISS) ? tri > 0 : tri >= 0);
        if (ok) {
            const crt_entity& e = s.entities[ent];
            const f3 lro = xform(e.inverse, ro, 1.0f);
            const f3 lrd = xform(e.inverse, rd, 0.0f);
            f3 a, b, c;
            if constexpr (FORMAT == CRT_FORMAT_WIDE4 || FORMAT == CRT_FORMAT_WIDE8) {
                const float4* rec = reinterpret_cast<const float4*>(s.tri_records) + 3 * int64_t(tri);
                const float4 ra = rec[0], rb = rec[1], rc = rec[2];
                a = mk(ra.x, ra.y, ra.z);
                b = mk(rb.x, rb.y, rb.z);
                c = mk(rc.x, rc.y, rc.z);
            } else {
                const int4 tr = *reinterpret_cast<const int4*>(s.tris[tri].packed);
                a = vertex_pos(s.verts, tr.x);
                b = vertex_pos(s.verts, tr.y);
                c = vertex_pos(s.verts, tr.z);
            }
            const f3 p = mk(lro.x + lrd.x * closest, lro.y + lrd.y * closest, lro.z + lrd.z * closest);
            const f3 uvw = barycentrics(p, a, b, c);
            tuvw = make_float4(closest, uvw.x, uvw.y, uvw.z);
        }
    }
    crt_hit h;
    h.t = tuvw.x;
    h.u = tuvw.y;
    h.v = tuvw.z;
    h.w = tuvw.w;
    h.tri = tri;
    h.mesh = mesh;
    h.entity = ent;
    h.status = pack_status(visits, status);
    tests = w.tests;
    return h;
}

template <int FORMAT, bool ANY, int LDS_DEPTH>
__device__ __forceinline__ void trace_one(const SceneArgs& s, const crt_ray& ray, uint32_t flags, int32_t max_iter,
                                          crt_hit* __restrict__ out, uint32_t* __restrict__ tri_tests,
                                          int32_t* lds, int32_t lds_stride) {
    uint32_t tests = 0;
    const crt_hit h = scene_trace<FORMAT, ANY, LDS_DEPTH>(s, ray, flags, max_iter, tests, lds, lds_stride);
    float4* o = reinterpret_cast<float4*>(out);
    o[0] = make_float4(h.t, h.u, h.v, h.w);
    o[1] = make_float4(__int_as_float(h.tri), __int_as_float(h.mesh), __int_as_float(h.entity),
                       __uint_as_float(h.status));
    if (tri_tests) *tri_tests = tests;
}

__device__ __forceinline__ crt_ray load_ray(const crt_ray* __restrict__ rays, int64_t i) {
    const float4* p = reinterpret_cast<const float4*>(rays + i);
    const float4 a = p[0], b = p[1];
    crt_ray r;
    r.org[0] = a.x;
    r.org[1] = a.y;
    r.org[2] = a.z;
    r.tmax = a.w;
    r.dir[0] = b.x;
    r.dir[1] = b.y;
    r.dir[2] = b.z;
    r.id = __float_as_uint(b.w);
    return r;
}

}  // namespace dev
}  // namespace crt
