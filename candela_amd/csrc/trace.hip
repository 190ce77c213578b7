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
// first half of its 32-byte record.  A wide node is eight 16-byte loads per
// group of four children and a triangle record three.  The stack walk keeps its first
// LDS_DEPTH entries in LDS (entry-major, so lane i of a wave touches bank i)
// and the rest in a per-lane scratch array.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"
#include "crt_internal.h"
#include "trace_launch.h"
#include "device_math.h"
#include "thread_walk.h"

#pragma clang fp contract(off)

namespace crt {
namespace dev {

using namespace ::crt::dm;

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
    // Stack format: one int per entry; wide formats: node + entry distance per
    // entry, then the pending leaves of one node.
    const size_t per_lane = FORMAT == CRT_FORMAT_STACK       ? size_t(D)
                            : FORMAT == CRT_FORMAT_STACKLESS ? 0
                            : FORMAT == CRT_FORMAT_WIDE8     ? size_t(wide::lds_ints<8, D>())
                                                             : size_t(wide::lds_ints<4, D>());
    const size_t lds = per_lane * block * sizeof(int32_t);
    hipLaunchKernelGGL((dev::trace_kernel<FORMAT, ANY, D>), dim3(unsigned(grid)), dim3(unsigned(block)), lds, stream,
                       s, rays, hits, n, flags, max_iter, tri_tests);
    return hipGetLastError();
}

template <int FORMAT, bool ANY>
static hipError_t launch_trace_f(const SceneArgs& s, const crt_ray* rays, crt_hit* hits, int64_t n, uint32_t flags,
                                 int32_t max_iter, int block, int depth, uint32_t* tri_tests, hipStream_t stream) {
    if constexpr (FORMAT == CRT_FORMAT_STACKLESS) {
        return launch_trace_d<FORMAT, ANY, 0>(s, rays, hits, n, flags, max_iter, block, tri_tests, stream);
    } else if constexpr (FORMAT == CRT_FORMAT_WIDE4 || FORMAT == CRT_FORMAT_WIDE8) {
        switch (depth) {
            case 8: return launch_trace_d<FORMAT, ANY, 8>(s, rays, hits, n, flags, max_iter, block, tri_tests, stream);
            case 32: return launch_trace_d<FORMAT, ANY, 32>(s, rays, hits, n, flags, max_iter, block, tri_tests, stream);
            default: return launch_trace_d<FORMAT, ANY, 16>(s, rays, hits, n, flags, max_iter, block, tri_tests, stream);
        }
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
    if (s.format == CRT_FORMAT_WIDE4) {
        return query == CRT_QUERY_ANY
                   ? launch_trace_f<CRT_FORMAT_WIDE4, true>(s, rays, hits, n, flags, max_iter, block, depth, tri_tests, stream)
                   : launch_trace_f<CRT_FORMAT_WIDE4, false>(s, rays, hits, n, flags, max_iter, block, depth, tri_tests, stream);
    }
    if (s.format == CRT_FORMAT_WIDE8) {
        return query == CRT_QUERY_ANY
                   ? launch_trace_f<CRT_FORMAT_WIDE8, true>(s, rays, hits, n, flags, max_iter, block, depth, tri_tests, stream)
                   : launch_trace_f<CRT_FORMAT_WIDE8, false>(s, rays, hits, n, flags, max_iter, block, depth, tri_tests, stream);
    }
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
