// Launchers of the wave-level stackless walk (wave_walk.h) for one ray batch:
//
//   grid        one ray per lane, ceil(n/256) workgroups; workgroups are dealt
//               to the 8 XCDs round robin, so blocks are remapped to give each
//               XCD (each with its own 4 MB L2) one contiguous eighth of the
//               batch -- neighbouring rays share the nodes near their origins;
//   persistent  device-filling waves that claim 64 rays at a time from a
//               counter until the batch is drained.
//
// Both produce hit records bit-identical to trace.hip and to the host oracle.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"
#include "trace_launch.h"
#include "wave_walk.h"

namespace crt {
namespace pdev {

template <bool ANY, int INNER, int LEAF_BATCH, int K>
__global__ void __launch_bounds__(256) grid_stackless(SceneArgs s, const crt_ray* __restrict__ rays,
                                                      crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                                      int32_t max_iter, uint32_t* __restrict__ tri_tests,
                                                      int32_t blocks) {
    const int32_t per = (blocks + 7) / 8;  // the grid is padded to a multiple of 8
    const int32_t logical = int32_t(blockIdx.x % 8u) * per + int32_t(blockIdx.x / 8u);
    if (logical >= blocks) return;
    const int64_t i = int64_t(logical) * 256 + threadIdx.x;
    WaveWalk<ANY, INNER, LEAF_BATCH, K> w;
    w.run(s, rays, hits, i < n ? i : -1, flags, max_iter, tri_tests);
}

template <bool ANY, int INNER, int LEAF_BATCH, int K>
__global__ void __launch_bounds__(256) persistent_stackless(SceneArgs s, const crt_ray* __restrict__ rays,
                                                            crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                                            int32_t max_iter,
                                                            unsigned long long* __restrict__ counter,
                                                            uint32_t* __restrict__ tri_tests) {
    WaveWalk<ANY, INNER, LEAF_BATCH, K> w;
    while (true) {
        unsigned long long base = 0;
        if (__lane_id() == 0) base = atomicAdd(counter, 64ull);
        base = (unsigned long long)__shfl((long long)base, 0);
        if (base >= (unsigned long long)n) break;
        const int64_t i = int64_t(base) + __lane_id();
        w.run(s, rays, hits, i < n ? i : -1, flags, max_iter, tri_tests);
    }
}

__global__ void reset_counter(unsigned long long* c) { *c = 0ull; }

}  // namespace pdev

// Presets (variant, INNER, LEAF_BATCH, K window nodes).
#define CRT_PRESETS(X) \
    X(0, 8, 8, 2)      \
    X(1, 8, 8, 1)      \
    X(2, 8, 8, 4)      \
    X(3, 16, 16, 2)    \
    X(4, 4, 4, 2)      \
    X(5, 8, 12, 4)

template <bool ANY>
static hipError_t launch_grid_any(const SceneArgs& s, const crt_ray* rays, crt_hit* hits, int64_t n, uint32_t flags,
                                  int32_t max_iter, uint32_t* tri_tests, int variant, hipStream_t stream) {
    const int32_t blocks = int32_t((n + 255) / 256);
    const unsigned padded = unsigned((blocks + 7) / 8 * 8);
#define CRT_GRID_CASE(V, I, LB, K)                                                                            \
    case V:                                                                                                  \
        hipLaunchKernelGGL((pdev::grid_stackless<ANY, I, LB, K>), dim3(padded), dim3(256), 0, stream, s, rays, \
                           hits, n, flags, max_iter, tri_tests, blocks);                                     \
        break;
    switch (variant) {
        CRT_PRESETS(CRT_GRID_CASE)
        default: return hipErrorInvalidValue;
    }
#undef CRT_GRID_CASE
    return hipGetLastError();
}

hipError_t launch_grid(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n, uint32_t flags,
                       int32_t max_iter, uint32_t* tri_tests, int variant, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    return query == CRT_QUERY_ANY
               ? launch_grid_any<true>(s, rays, hits, n, flags, max_iter, tri_tests, variant, stream)
               : launch_grid_any<false>(s, rays, hits, n, flags, max_iter, tri_tests, variant, stream);
}

template <bool ANY>
static hipError_t launch_p(const SceneArgs& s, const crt_ray* rays, crt_hit* hits, int64_t n, uint32_t flags,
                           int32_t max_iter, unsigned long long* counter, uint32_t* tri_tests, int grid,
                           int variant, hipStream_t stream) {
#define CRT_P_CASE(V, I, LB, K)                                                                                    \
    case V:                                                                                                       \
        hipLaunchKernelGGL((pdev::persistent_stackless<ANY, I, LB, K>), dim3(unsigned(grid)), dim3(256), 0, stream, \
                           s, rays, hits, n, flags, max_iter, counter, tri_tests);                                \
        break;
    switch (variant) {
        CRT_PRESETS(CRT_P_CASE)
        default: return hipErrorInvalidValue;
    }
#undef CRT_P_CASE
    return hipGetLastError();
}

int persistent_fill_grid(int query) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    const void* fn = query == CRT_QUERY_ANY
                         ? reinterpret_cast<const void*>(&pdev::persistent_stackless<true, 8, 8, 2>)
                         : reinterpret_cast<const void*>(&pdev::persistent_stackless<false, 8, 8, 2>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess || per_cu <= 0) per_cu = 4;
    return cus * per_cu;
}

hipError_t launch_persistent(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                             uint32_t flags, int32_t max_iter, unsigned long long* counter, uint32_t* tri_tests,
                             const PersistentShape& shape, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(pdev::reset_counter, dim3(1), dim3(1), 0, stream, counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t need = (n + 255) / 256;
    const int grid = int(need < shape.grid ? need : shape.grid);
    return query == CRT_QUERY_ANY
               ? launch_p<true>(s, rays, hits, n, flags, max_iter, counter, tri_tests, grid, shape.variant, stream)
               : launch_p<false>(s, rays, hits, n, flags, max_iter, counter, tri_tests, grid, shape.variant, stream);
}

}  // namespace crt
