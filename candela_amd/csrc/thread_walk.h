// One ray per lane: the reference's walks and scene loop as plain device
// functions (used by the one-ray-per-lane kernels of trace.hip and inline by
// the shading kernels of shade.hip).  Each function restates one GLSL
// function of Source/Core/Shaders/Intersectors/Include/:
//   walk_stackless<ANY>  TraverseBVHStackless.glsl:175-278 / :463-556
//   walk_stack<ANY>      TraverseBVHStack.glsl:168-324 / :509-657
//   walk_wide<ANY, W>    the same closest / any-hit contract over the wide
//                        layout derived from either (no reference counterpart)
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
                lds[sp * lds_stride] = post;
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
        cur = sp < LDS_DEPTH ? lds[sp * lds_stride] : spill[sp - LDS_DEPTH];
    }
    if (!stopped) status |= CRT_STATUS_ITER_CAP;
    return uint32_t(it);
}

// Triangles [start, start+len) of the wide layout's records (three 16-byte
// loads each: v0 + triangle index, v1 + mesh id, v2).  Same test and
// acceptance as test_leaf.
template <bool ANY>
__device__ __forceinline__ bool test_records(const SceneArgs& s, f3 ro, f3 rd, int32_t start, int32_t len,
                                             float& tmax, Walk& w) {
    const float4* __restrict__ rec = reinterpret_cast<const float4*>(s.tri_records);
    for (int32_t idx = start; idx < start + len; ++idx) {
        const float4 a = rec[3 * int64_t(idx)], b = rec[3 * int64_t(idx) + 1], c = rec[3 * int64_t(idx) + 2];
        const float t = tri_t(ro, rd, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), mk(c.x, c.y, c.z));
        w.tests++;
        if (t > 0.0f && t < tmax) {
            tmax = t;
            w.t = t;
            w.tri = __float_as_int(a.w);
            w.mesh = __float_as_int(b.w);
            if (ANY) return true;
        }
    }
    return false;
}

// Relative margin of the wide walk's box tests.  A triangle lying on a face
// of its leaf box (touching boxes, axis-aligned geometry) can have a computed
// t a few ulps below the box's computed entry distance; with a closer-by-ulps
// hit already found, an exact comparison would cull the leaf holding the true
// nearest triangle.  Growing the exit distance and tmax by 2^-19 (16 ulps)
// keeps such leaves, so the walk reports the brute-force nearest hit.
constexpr float kWideGrow = 1.0f + 1.0f / 524288.0f;

// Walk of a wide tree (CRT_FORMAT_WIDE4 / WIDE8, W children per node, W / 4
// groups of 128 bytes).  Per node: every non-empty child box is tested with
// the slab arithmetic of the stack include's RayBounds (entry t0 =
// max(near, 0), exit t1 = min(far, tmax), both upper ends grown by
// kWideGrow; hit when t1 >= t0 and t0 or t1 is positive); hit leaves are
// tested at once; of the hit inner children the nearest is visited next and
// the others are pushed with their entry distance.  A popped entry whose
// entry distance exceeds the grown current closest t is skipped -- exactly
// the entries whose box test would now fail.  Stack entries (node, entry
// distance) live in LDS for the first LDS_DEPTH, then in per-lane scratch;
// beyond LDS_DEPTH + 64 the walk stops with STACK_OVERFLOW.
template <bool ANY, int W, int LDS_DEPTH>
__device__ __forceinline__ uint32_t walk_wide(const SceneArgs& s, f3 ro, f3 rd, int32_t root, int32_t count,
                                              float tmax, int32_t max_iter, Walk& w, uint32_t& status, int32_t* lds,
                                              int32_t lds_stride) {
    constexpr int G = W / 4;
    constexpr int kSpill = 64;
    const float4* __restrict__ groups = reinterpret_cast<const float4*>(s.nodes);
    const f3 inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    int32_t spill_node[kSpill];
    float spill_t[kSpill];
    int32_t sp = 0;
    int32_t it = 0;
    int32_t cur = root;
    bool stopped = false;
    w.t = -1.0f;
    w.tri = -1;
    w.mesh = -1;
    while (it < max_iter) {
        if (cur < root || cur >= root + count || int64_t(cur) >= s.total_nodes) {
            status |= CRT_STATUS_BAD_INDEX;
            stopped = true;
            break;
        }
        it++;
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
                if (cn[k] < 0) continue;
                const float fx = (hx[k] - ro.x) * inv.x, fy = (hy[k] - ro.y) * inv.y, fz = (hz[k] - ro.z) * inv.z;
                const float nx = (lx[k] - ro.x) * inv.x, ny = (ly[k] - ro.y) * inv.y, nz = (lz[k] - ro.z) * inv.z;
                const float t1 = mn(mn(mx(fx, nx), mn(mx(fy, ny), mx(fz, nz))) * kWideGrow, tmax * kWideGrow);
                const float t0 = mx(mx(mn(fx, nx), mx(mn(fy, ny), mn(fz, nz))), 0.0f);
                if (!(t1 >= t0) || !(t0 > 0.0f || t1 > 0.0f)) continue;
                if (cn[k] > 0) {
                    if (test_records<ANY>(s, ro, rd, ch[k], cn[k], tmax, w)) return uint32_t(it);
                    continue;
                }
                int32_t push_node = ch[k];
                float push_t = t0;
                if (near_node < 0) {
                    near_node = ch[k];
                    near_t = t0;
                    continue;
                }
                if (t0 < near_t) {
                    push_node = near_node;
                    push_t = near_t;
                    near_node = ch[k];
                    near_t = t0;
                }
                if (sp >= LDS_DEPTH + kSpill) {
                    overflow = true;
                    continue;
                }
                if (sp < LDS_DEPTH) {
                    lds[(2 * sp) * lds_stride] = push_node;
                    lds[(2 * sp + 1) * lds_stride] = __float_as_int(push_t);
                } else {
                    spill_node[sp - LDS_DEPTH] = push_node;
                    spill_t[sp - LDS_DEPTH] = push_t;
                }
                ++sp;
            }
        }
        if (overflow) {
            status |= CRT_STATUS_STACK_OVERFLOW;
            stopped = true;
            break;
        }
        if (near_node >= 0 && (ANY || !(near_t > tmax * kWideGrow))) {
            cur = near_node;
            continue;
        }
        bool found = false;
        while (sp > 0) {
            --sp;
            int32_t node;
            float t;
            if (sp < LDS_DEPTH) {
                node = lds[(2 * sp) * lds_stride];
                t = __int_as_float(lds[(2 * sp + 1) * lds_stride]);
            } else {
                node = spill_node[sp - LDS_DEPTH];
                t = spill_t[sp - LDS_DEPTH];
            }
            if (!ANY && t > tmax * kWideGrow) continue;
            cur = node;
            found = true;
            break;
        }
        if (!found) {
            stopped = true;
            break;
        }
    }
    if (!stopped) status |= CRT_STATUS_ITER_CAP;
    return uint32_t(it);
}

template <int FORMAT, bool ANY, int LDS_DEPTH>
__device__ __forceinline__ uint32_t walk_entity(const SceneArgs& s, f3 ro, f3 rd, int32_t start, int32_t count,
                                                float tmax, int32_t max_iter, Walk& w, uint32_t& status,
                                                int32_t* lds, int32_t lds_stride) {
    if constexpr (FORMAT == CRT_FORMAT_STACKLESS)
        return walk_stackless<ANY>(s, ro, rd, start, count, tmax, max_iter, w, status);
    else if constexpr (FORMAT == CRT_FORMAT_STACK)
        return walk_stack<ANY, LDS_DEPTH>(s, ro, rd, start, count, tmax, max_iter, w, status, lds, lds_stride);
    else
        return walk_wide<ANY, FORMAT == CRT_FORMAT_WIDE8 ? 8 : 4, LDS_DEPTH>(s, ro, rd, start, count, tmax,
                                                                              max_iter, w, status, lds, lds_stride);
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
                                                           max_iter, w, status, lds, lds_stride);
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
                                                            max_iter, w, status, lds, lds_stride);
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
