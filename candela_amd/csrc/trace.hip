// HIP traversal kernels for gfx950 (MI355X) -- the Ray Intersection API.
//
// One ray per lane, 64-lane waves.  Each kernel restates one GLSL function of
// the reference (paths relative to Source/Core/Shaders/Intersectors/):
//   walk_stackless<ANY>  Include/TraverseBVHStackless.glsl:175-278 / :463-556
//   walk_stack<ANY>      Include/TraverseBVHStack.glsl:168-324 / :509-657
//   scene loop           Include/TraverseBVHStackless.glsl:280-366, :558-579
//   primary ray          TraverseBVHStackless.glsl:143-148, :371-388
//
// Arithmetic is FP32 with FMA contraction off and IEEE division / square
// root, in the same operation order as the host restatement, so t, u, v, w
// and the hit triangle are bit-identical to it.  min/max use IEEE minNum
// (v_min_f32 / v_max_f32), which the GLSL leaves open for NaN operands.
//
// Memory: a stackless node is two 16-byte loads, a stack node four, a
// triangle one 16-byte load and each vertex position one 16-byte load of the
// first half of its 32-byte record.  The stack walk keeps its first
// LDS_DEPTH entries in LDS (entry-major, so lane i of a wave touches bank i)
// and the rest in a per-lane scratch array.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"
#include "crt_internal.h"
#include "trace_launch.h"
#include "device_math.h"

#pragma clang fp contract(off)

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

template <int FORMAT, bool ANY, int LDS_DEPTH>
__device__ __forceinline__ uint32_t walk_entity(const SceneArgs& s, f3 ro, f3 rd, int32_t start, int32_t count,
                                                float tmax, int32_t max_iter, Walk& w, uint32_t& status,
                                                int32_t* lds, int32_t lds_stride) {
    if constexpr (FORMAT == CRT_FORMAT_STACKLESS)
        return walk_stackless<ANY>(s, ro, rd, start, count, tmax, max_iter, w, status);
    else
        return walk_stack<ANY, LDS_DEPTH>(s, ro, rd, start, count, tmax, max_iter, w, status, lds, lds_stride);
}


// The per-ray scene loop: IntersectScene / IntersectSceneIgnoreTransparent
// (closest) or the occlusion IntersectScene (any).
template <int FORMAT, bool ANY, int LDS_DEPTH>
__device__ __forceinline__ void trace_one(const SceneArgs& s, const crt_ray& ray, uint32_t flags, int32_t max_iter,
                                          crt_hit* __restrict__ out, uint32_t* __restrict__ tri_tests,
                                          int32_t* lds, int32_t lds_stride) {
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
            const int4 tr = *reinterpret_cast<const int4*>(s.tris[tri].packed);
            const f3 a = vertex_pos(s.verts, tr.x), b = vertex_pos(s.verts, tr.y), c = vertex_pos(s.verts, tr.z);
            const f3 p = mk(lro.x + lrd.x * closest, lro.y + lrd.y * closest, lro.z + lrd.z * closest);
            const f3 uvw = barycentrics(p, a, b, c);
            tuvw = make_float4(closest, uvw.x, uvw.y, uvw.z);
        }
    }
    float4* o = reinterpret_cast<float4*>(out);
    o[0] = tuvw;
    o[1] = make_float4(__int_as_float(tri), __int_as_float(mesh), __int_as_float(ent),
                       __uint_as_float(pack_status(visits, status)));
    if (tri_tests) *tri_tests = w.tests;
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

template <int FORMAT, bool ANY, int LDS_DEPTH>
__global__ void __launch_bounds__(1024) trace_kernel(SceneArgs s, const crt_ray* __restrict__ rays,
                                                     crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                                     int32_t max_iter, uint32_t* __restrict__ tri_tests) {
    extern __shared__ int32_t lds_stack[];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const crt_ray r = load_ray(rays, i);
    trace_one<FORMAT, ANY, LDS_DEPTH>(s, r, flags, max_iter, hits + i, tri_tests ? tri_tests + i : nullptr,
                                      lds_stack + threadIdx.x, int32_t(blockDim.x));
}

// IntersectPrimary: ray of pixel (x, y) as in TraverseBVHStackless.glsl:371-378
// (uv = pixel / dims, GetRayDirectionAt, normalize), then closest hit.
template <int FORMAT, int LDS_DEPTH>
__global__ void __launch_bounds__(1024) primary_kernel(SceneArgs s, Camera cam, crt_ray* __restrict__ rays,
                                                       crt_hit* __restrict__ hits, int32_t width, int32_t height,
                                                       uint32_t flags, int32_t max_iter) {
    extern __shared__ int32_t lds_stack[];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= int64_t(width) * height) return;
    const int32_t x = int32_t(i % width), y = int32_t(i / width);
    const float u = float(x) / float(width), v = float(y) / float(height);
    const float cx = u * 2.0f - 1.0f, cy = v * 2.0f - 1.0f;
    const float* ip = cam.inv_proj;
    const float ex = ip[0] * cx + ip[4] * cy + ip[8] * -1.0f + ip[12] * 1.0f;
    const float ey = ip[1] * cx + ip[5] * cy + ip[9] * -1.0f + ip[13] * 1.0f;
    const f3 d = xform(cam.inv_view, mk(ex, ey, -1.0f), 0.0f);
    const float il = 1.0f / sqrtf(dot(d, d));
    crt_ray r;
    r.org[0] = cam.inv_view[12];
    r.org[1] = cam.inv_view[13];
    r.org[2] = cam.inv_view[14];
    r.tmax = cam.tmax;
    r.dir[0] = d.x * il;
    r.dir[1] = d.y * il;
    r.dir[2] = d.z * il;
    r.id = uint32_t(i);
    float4* rp = reinterpret_cast<float4*>(rays + i);
    rp[0] = make_float4(r.org[0], r.org[1], r.org[2], r.tmax);
    rp[1] = make_float4(r.dir[0], r.dir[1], r.dir[2], __uint_as_float(r.id));
    trace_one<FORMAT, false, LDS_DEPTH>(s, r, flags, max_iter, hits + i, nullptr, lds_stack + threadIdx.x,
                                        int32_t(blockDim.x));
}

}  // namespace dev

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

template <int FORMAT, bool ANY, int D>
static hipError_t launch_trace_d(const SceneArgs& s, const crt_ray* rays, crt_hit* hits, int64_t n, uint32_t flags,
                                 int32_t max_iter, int block, uint32_t* tri_tests, hipStream_t stream) {
    const int64_t grid = (n + block - 1) / block;
    const size_t lds = FORMAT == CRT_FORMAT_STACK ? size_t(D) * block * sizeof(int32_t) : 0;
    hipLaunchKernelGGL((dev::trace_kernel<FORMAT, ANY, D>), dim3(unsigned(grid)), dim3(unsigned(block)), lds, stream,
                       s, rays, hits, n, flags, max_iter, tri_tests);
    return hipGetLastError();
}

template <int FORMAT, bool ANY>
static hipError_t launch_trace_f(const SceneArgs& s, const crt_ray* rays, crt_hit* hits, int64_t n, uint32_t flags,
                                 int32_t max_iter, int block, int depth, uint32_t* tri_tests, hipStream_t stream) {
    if constexpr (FORMAT == CRT_FORMAT_STACKLESS) {
        return launch_trace_d<FORMAT, ANY, 0>(s, rays, hits, n, flags, max_iter, block, tri_tests, stream);
    } else {
        switch (depth) {
            case 8: return launch_trace_d<FORMAT, ANY, 8>(s, rays, hits, n, flags, max_iter, block, tri_tests, stream);
            case 16: return launch_trace_d<FORMAT, ANY, 16>(s, rays, hits, n, flags, max_iter, block, tri_tests, stream);
            case 32: return launch_trace_d<FORMAT, ANY, 32>(s, rays, hits, n, flags, max_iter, block, tri_tests, stream);
            case 48: return launch_trace_d<FORMAT, ANY, 48>(s, rays, hits, n, flags, max_iter, block, tri_tests, stream);
            default: return launch_trace_d<FORMAT, ANY, 64>(s, rays, hits, n, flags, max_iter, block, tri_tests, stream);
        }
    }
}

hipError_t launch_trace(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                        uint32_t flags, int32_t max_iter, int block, int depth, uint32_t* tri_tests,
                        hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    if (s.format == CRT_FORMAT_STACKLESS) {
        return query == CRT_QUERY_ANY
                   ? launch_trace_f<CRT_FORMAT_STACKLESS, true>(s, rays, hits, n, flags, max_iter, block, depth, tri_tests, stream)
                   : launch_trace_f<CRT_FORMAT_STACKLESS, false>(s, rays, hits, n, flags, max_iter, block, depth,
                                                                 tri_tests, stream);
    }
    return query == CRT_QUERY_ANY
               ? launch_trace_f<CRT_FORMAT_STACK, true>(s, rays, hits, n, flags, max_iter, block, depth, tri_tests, stream)
               : launch_trace_f<CRT_FORMAT_STACK, false>(s, rays, hits, n, flags, max_iter, block, depth, tri_tests, stream);
}

hipError_t launch_primary(const SceneArgs& s, const Camera& cam, crt_ray* rays, crt_hit* hits, int width, int height,
                          uint32_t flags, int32_t max_iter, int block, hipStream_t stream) {
    const int64_t n = int64_t(width) * height;
    if (n <= 0) return hipSuccess;
    const int64_t grid = (n + block - 1) / block;
    if (s.format == CRT_FORMAT_STACKLESS) {
        hipLaunchKernelGGL((dev::primary_kernel<CRT_FORMAT_STACKLESS, 0>), dim3(unsigned(grid)), dim3(unsigned(block)),
                           0, stream, s, cam, rays, hits, width, height, flags, max_iter);
    } else {
        const size_t lds = size_t(16) * block * sizeof(int32_t);
        hipLaunchKernelGGL((dev::primary_kernel<CRT_FORMAT_STACK, 16>), dim3(unsigned(grid)), dim3(unsigned(block)),
                           lds, stream, s, cam, rays, hits, width, height, flags, max_iter);
    }
    return hipGetLastError();
}

}  // namespace crt

// ---------------------------------------------------------------------------
// Streaming copy (bandwidth / counter calibration)
// ---------------------------------------------------------------------------
namespace crt {
namespace dev {
__global__ void __launch_bounds__(256) copy16_kernel(float4* __restrict__ dst, const float4* __restrict__ src,
                                                     int64_t n) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}
}  // namespace dev

hipError_t launch_copy16(void* dst, const void* src, int64_t n16, hipStream_t stream) {
    if (n16 <= 0) return hipSuccess;
    const int64_t want = (n16 + 255) / 256;
    const unsigned grid = unsigned(want < 256 * 64 ? want : 256 * 64);
    hipLaunchKernelGGL(dev::copy16_kernel, dim3(grid), dim3(256), 0, stream, static_cast<float4*>(dst),
                       static_cast<const float4*>(src), n16);
    return hipGetLastError();
}
}  // namespace crt
