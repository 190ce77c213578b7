// Launchers of the speculative walks of the reference's two BVH2 layouts
// (spec_bvh2.h), crt_trace kernels 4 (grid) and 5 (refilling) over
// CRT_FORMAT_STACKLESS (sbvh2::grid / persistent with D = 0) and
// CRT_FORMAT_STACK (D > 0 stack entries in LDS):
//
//   grid        64 consecutive rays per wave, ceil(n/256) workgroups in the
//               XCD-aware block order of the other grid kernels; the records
//               are deterministic (a host twin replays the schedule);
//   persistent  device-filling waves that refill idle lanes from the sharded
//               queue (queue.h).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"
#include "crt_internal.h"
#include "spec_bvh2.h"
#include "trace_launch.h"

#pragma clang fp contract(off)

namespace crt {
namespace sbvh2 {

// Per workgroup: Q queue slots per lane, D stack entries per lane (stack
// layout), the leaf-phase list and keys of each wave.
template <int D, int Q>
struct GroupLds {
    int32_t queue[Q * 256];
    int32_t stack[D * 256];
    spec::WaveLds<Q> waves[4];
};

template <int Q>
struct GroupLds<0, Q> {
    int32_t queue[Q * 256];
    spec::WaveLds<Q> waves[4];
};

template <bool ANY, int D, int I, int Q>
__device__ __forceinline__ void bind(SpecBvh2Wave<ANY, D, I, Q>& w, GroupLds<D, Q>& lds, int32_t* spill) {
    w.q = (LdsInt*)(lds.queue + threadIdx.x);
    if constexpr (D > 0) w.st = sdev::NodeStack<D>{lds.stack + threadIdx.x, 256, spill};
}

template <bool ANY, int D, int I, int Q, int OCC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1, 8)))
grid(SceneArgs s, const crt_ray* __restrict__ rays, crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
     int32_t max_iter, uint32_t* __restrict__ tri_tests, int32_t blocks) {
    __shared__ GroupLds<D, Q> lds;
    const int32_t per = (blocks + 7) / 8;  // the grid is padded to a multiple of 8
    const int32_t logical = int32_t(blockIdx.x % 8u) * per + int32_t(blockIdx.x / 8u);
    if (logical >= blocks) return;
    const int64_t i = int64_t(logical) * 256 + threadIdx.x;
    int32_t spill[D > 0 ? sdev::spill_entries<D>() : 1];
    SpecBvh2Wave<ANY, D, I, Q> w;
    bind(w, lds, spill);
    w.run(s, rays, hits, i < n ? i : -1, flags, max_iter, tri_tests, lds.waves[threadIdx.x >> 6]);
}

template <bool ANY, int D, int I, int Q, int REFILL, int OCC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1, 8)))
persistent(SceneArgs s, const crt_ray* __restrict__ rays, crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
           int32_t max_iter, uint32_t* __restrict__ tri_tests, unsigned long long* __restrict__ counters) {
    __shared__ GroupLds<D, Q> lds;
    int32_t spill[D > 0 ? sdev::spill_entries<D>() : 1];
    SpecBvh2Wave<ANY, D, I, Q> w;
    bind(w, lds, spill);
    w.run_refill(s, rays, hits, n, flags, max_iter, tri_tests, lds.waves[threadIdx.x >> 6], counters, REFILL);
}

}  // namespace sbvh2

// Presets of the grid walks (variant, INNER node steps per round, Q queue
// slots per lane, OCC waves per SIMD; 0 = the compiler's choice).  The stack
// layout keeps 8 stack entries per lane in LDS.
#define CRT_SBVH2_PRESETS(X) \
    X(0, 8, 8, 0)            \
    X(1, 16, 8, 0)           \
    X(2, 16, 12, 0)          \
    X(3, 32, 16, 0)

// Presets of the refilling walks (variant, INNER, Q, REFILL idle lanes, OCC).
#define CRT_SBVH2_P_PRESETS(X) \
    X(0, 16, 8, 16, 0)         \
    X(1, 32, 16, 16, 0)        \
    X(2, 16, 12, 8, 0)         \
    X(3, 32, 16, 32, 0)

constexpr int kStackLds = 8;

template <bool ANY, int D>
static const void* sbvh2_grid_fn(int variant) {
#define CRT_SBVH2_G_FN(V, I, Q, OCC) \
    case V: return reinterpret_cast<const void*>(&sbvh2::grid<ANY, D, I, Q, OCC>);
    switch (variant) {
        CRT_SBVH2_PRESETS(CRT_SBVH2_G_FN)
        default: return nullptr;
    }
#undef CRT_SBVH2_G_FN
}

template <bool ANY, int D>
static const void* sbvh2_persistent_fn(int variant) {
#define CRT_SBVH2_P_FN(V, I, Q, R, OCC) \
    case V: return reinterpret_cast<const void*>(&sbvh2::persistent<ANY, D, I, Q, R, OCC>);
    switch (variant) {
        CRT_SBVH2_P_PRESETS(CRT_SBVH2_P_FN)
        default: return nullptr;
    }
#undef CRT_SBVH2_P_FN
}

static const void* grid_fn(bool stack, int query, int variant) {
    const bool any = query == CRT_QUERY_ANY;
    if (stack) return any ? sbvh2_grid_fn<true, kStackLds>(variant) : sbvh2_grid_fn<false, kStackLds>(variant);
    return any ? sbvh2_grid_fn<true, 0>(variant) : sbvh2_grid_fn<false, 0>(variant);
}

static const void* persistent_fn(bool stack, int query, int variant) {
    const bool any = query == CRT_QUERY_ANY;
    if (stack)
        return any ? sbvh2_persistent_fn<true, kStackLds>(variant) : sbvh2_persistent_fn<false, kStackLds>(variant);
    return any ? sbvh2_persistent_fn<true, 0>(variant) : sbvh2_persistent_fn<false, 0>(variant);
}

hipError_t launch_spec_bvh2_grid(const SceneArgs& s, bool stack, int query, const crt_ray* rays, crt_hit* hits,
                                 int64_t n, uint32_t flags, int32_t max_iter, uint32_t* tri_tests, int variant,
                                 hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const void* fn = grid_fn(stack, query, variant);
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

hipError_t launch_spec_bvh2_persistent(const SceneArgs& s, bool stack, int query, const crt_ray* rays, crt_hit* hits,
                                       int64_t n, uint32_t flags, int32_t max_iter, unsigned long long* counters,
                                       uint32_t* tri_tests, int variant, int grid, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const void* fn = persistent_fn(stack, query, variant);
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
