// Launchers of the wave-scheduled walk of the reference's stack layout
// (stack_walk.h) for one ray batch:
//
//   grid        one ray per lane, ceil(n/256) workgroups, XCD-aware block order
//               (consecutive ray tiles on one XCD, whose L2 then holds the part
//               of the tree those rays share);
//   persistent  device-filling waves that claim one ray per idle lane from the
//               sharded queue (queue.h) once REFILL lanes are idle.
//
// Both give hit records bit-identical to trace.hip's one-ray-per-lane stack
// walk and to the host oracle.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"
#include "crt_internal.h"
#include "stack_walk.h"
#include "trace_launch.h"

#pragma clang fp contract(off)

namespace crt {
namespace sdev {

template <bool ANY, int D, int I, int LB>
__device__ __forceinline__ void grid_body(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                          crt_hit* __restrict__ hits, int64_t n, uint32_t flags, int32_t max_iter,
                                          uint32_t* __restrict__ tri_tests, int32_t blocks, int32_t* lds) {
    const int32_t per = (blocks + 7) / 8;  // the grid is padded to a multiple of 8
    const int32_t logical = int32_t(blockIdx.x % 8u) * per + int32_t(blockIdx.x / 8u);
    if (logical >= blocks) return;
    const int64_t i = int64_t(logical) * 256 + threadIdx.x;
    int32_t spill[spill_entries<D>()];
    NodeStack<D> st{lds + threadIdx.x, 256, spill};
    StackWave<ANY, D, I, LB> w;
    w.run(s, rays, hits, i < n ? i : -1, flags, max_iter, tri_tests, st);
}

template <bool ANY, int D, int I, int LB>
__global__ void __launch_bounds__(256) stack_grid(SceneArgs s, const crt_ray* __restrict__ rays,
                                                  crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                                  int32_t max_iter, uint32_t* __restrict__ tri_tests, int32_t blocks) {
    __shared__ int32_t lds[D * 256];
    grid_body<ANY, D, I, LB>(s, rays, hits, n, flags, max_iter, tri_tests, blocks, lds);
}

// The same kernel compiled for at least OCC waves per SIMD.
template <bool ANY, int D, int I, int LB, int OCC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC, 8)))
stack_grid_occ(SceneArgs s, const crt_ray* __restrict__ rays, crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
               int32_t max_iter, uint32_t* __restrict__ tri_tests, int32_t blocks) {
    __shared__ int32_t lds[D * 256];
    grid_body<ANY, D, I, LB>(s, rays, hits, n, flags, max_iter, tri_tests, blocks, lds);
}

template <bool ANY, int D, int I, int LB, int R>
__device__ __forceinline__ void persistent_body(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                                crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                                int32_t max_iter, uint32_t* __restrict__ tri_tests,
                                                unsigned long long* __restrict__ counters, int32_t* lds) {
    int32_t spill[spill_entries<D>()];
    NodeStack<D> st{lds + threadIdx.x, 256, spill};
    StackWave<ANY, D, I, LB> w;
    w.run_refill(s, rays, hits, n, flags, max_iter, tri_tests, st, counters, R);
}

template <bool ANY, int D, int I, int LB, int R>
__global__ void __launch_bounds__(256) stack_persistent(SceneArgs s, const crt_ray* __restrict__ rays,
                                                        crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                                        int32_t max_iter, uint32_t* __restrict__ tri_tests,
                                                        unsigned long long* __restrict__ counters) {
    __shared__ int32_t lds[D * 256];
    persistent_body<ANY, D, I, LB, R>(s, rays, hits, n, flags, max_iter, tri_tests, counters, lds);
}

template <bool ANY, int D, int I, int LB, int R, int OCC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC, 8)))
stack_persistent_occ(SceneArgs s, const crt_ray* __restrict__ rays, crt_hit* __restrict__ hits, int64_t n,
                     uint32_t flags, int32_t max_iter, uint32_t* __restrict__ tri_tests,
                     unsigned long long* __restrict__ counters) {
    __shared__ int32_t lds[D * 256];
    persistent_body<ANY, D, I, LB, R>(s, rays, hits, n, flags, max_iter, tri_tests, counters, lds);
}

template <bool ANY, int D, int I, int LB, int OCC>
const void* grid_kernel() {
    if constexpr (OCC == 0)
        return reinterpret_cast<const void*>(&stack_grid<ANY, D, I, LB>);
    else
        return reinterpret_cast<const void*>(&stack_grid_occ<ANY, D, I, LB, OCC>);
}

template <bool ANY, int D, int I, int LB, int R, int OCC>
const void* persistent_kernel() {
    if constexpr (OCC == 0)
        return reinterpret_cast<const void*>(&stack_persistent<ANY, D, I, LB, R>);
    else
        return reinterpret_cast<const void*>(&stack_persistent_occ<ANY, D, I, LB, R, OCC>);
}

}  // namespace sdev

// Presets of the grid walk (variant, LDS stack entries, INNER node steps per
// round, LEAF_BATCH parked lanes, OCC waves per SIMD; 0 = the compiler's
// choice).
#define CRT_STACK_PRESETS(X) \
    X(0, 16, 4, 16, 0)       \
    X(1, 16, 8, 16, 0)       \
    X(2, 8, 4, 16, 0)        \
    X(3, 16, 4, 32, 0)       \
    X(4, 16, 2, 8, 0)        \
    X(5, 8, 8, 32, 0)        \
    X(6, 16, 4, 16, 6)       \
    X(7, 8, 4, 16, 6)

// Presets of the refilling walk (variant, LDS stack entries, INNER,
// LEAF_BATCH, REFILL idle lanes, OCC).
#define CRT_STACK_P_PRESETS(X) \
    X(0, 16, 4, 16, 16, 0)     \
    X(1, 16, 4, 32, 16, 0)     \
    X(2, 16, 4, 16, 8, 0)      \
    X(3, 8, 4, 16, 16, 0)      \
    X(4, 16, 8, 32, 24, 0)     \
    X(5, 16, 2, 16, 16, 0)     \
    X(6, 16, 4, 16, 16, 6)     \
    X(7, 8, 4, 16, 8, 6)

template <bool ANY>
static const void* stack_grid_fn(int variant) {
#define CRT_STACK_G_FN(V, D, I, LB, OCC) \
    case V: return sdev::grid_kernel<ANY, D, I, LB, OCC>();
    switch (variant) {
        CRT_STACK_PRESETS(CRT_STACK_G_FN)
        default: return nullptr;
    }
#undef CRT_STACK_G_FN
}

template <bool ANY>
static const void* stack_persistent_fn(int variant) {
#define CRT_STACK_P_FN(V, D, I, LB, R, OCC) \
    case V: return sdev::persistent_kernel<ANY, D, I, LB, R, OCC>();
    switch (variant) {
        CRT_STACK_P_PRESETS(CRT_STACK_P_FN)
        default: return nullptr;
    }
#undef CRT_STACK_P_FN
}

hipError_t launch_stack_grid(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                             uint32_t flags, int32_t max_iter, uint32_t* tri_tests, int variant, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const void* fn = query == CRT_QUERY_ANY ? stack_grid_fn<true>(variant) : stack_grid_fn<false>(variant);
    if (!fn) return hipErrorInvalidValue;
    int32_t blocks = int32_t((n + 255) / 256);
    const unsigned padded = unsigned((blocks + 7) / 8 * 8);
    SceneArgs sa = s;
    const crt_ray* r = rays;
    crt_hit* h = hits;
    int64_t nn = n;
    uint32_t fl = flags;
    int32_t mi = max_iter;
    uint32_t* tt = tri_tests;
    void* args[] = {&sa, &r, &h, &nn, &fl, &mi, &tt, &blocks};
    const hipError_t e = hipLaunchKernel(fn, dim3(padded), dim3(256), args, 0, stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t launch_stack_persistent(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                                   uint32_t flags, int32_t max_iter, unsigned long long* counters,
                                   uint32_t* tri_tests, int variant, int grid, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const void* fn = query == CRT_QUERY_ANY ? stack_persistent_fn<true>(variant) : stack_persistent_fn<false>(variant);
    if (!fn) return hipErrorInvalidValue;
    if (grid <= 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess || per_cu <= 0) per_cu = 4;
        grid = cus * per_cu;
    }
    const int64_t need = (n + 255) / 256;
    if (need < grid) grid = int(need);
    hipLaunchKernelGGL(reset_queue, dim3(1), dim3(64), 0, stream, counters);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    SceneArgs sa = s;
    const crt_ray* r = rays;
    crt_hit* h = hits;
    int64_t nn = n;
    uint32_t fl = flags;
    int32_t mi = max_iter;
    uint32_t* tt = tri_tests;
    unsigned long long* c = counters;
    void* args[] = {&sa, &r, &h, &nn, &fl, &mi, &tt, &c};
    e = hipLaunchKernel(fn, dim3(unsigned(grid)), dim3(256), args, 0, stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

}  // namespace crt
