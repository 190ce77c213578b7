// Wave-scheduled walk of the wide layout (BVH4 / BVH8): one ray per lane,
// leaves parked and tested by the wave together (wide_walk.h: WideWave), and
// the XCD-aware block order of the stackless grid kernel: blocks that trace
// consecutive ray tiles run on one XCD, whose L2 then holds the part of the
// tree those rays share.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"
#include "crt_internal.h"
#include "trace_launch.h"
#include "wide_walk.h"

#pragma clang fp contract(off)

namespace crt {

static_assert(kQueueWords <= kQueueBlockWords, "a queue block holds every shard counter");
namespace wdev {

template <bool ANY, int W, int LDS_DEPTH, int INNER, int LEAF_BATCH>
__device__ __forceinline__ void grid_body(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                          crt_hit* __restrict__ hits, int64_t n, uint32_t flags, int32_t max_iter,
                                          uint32_t* __restrict__ tri_tests, int32_t blocks, int32_t* lds) {
    const int32_t per = (blocks + 7) / 8;  // the grid is padded to a multiple of 8
    const int32_t logical = int32_t(blockIdx.x % 8u) * per + int32_t(blockIdx.x / 8u);
    if (logical >= blocks) return;
    const int64_t i = int64_t(logical) * 256 + threadIdx.x;
    wide::WideWave<ANY, W, LDS_DEPTH, INNER, LEAF_BATCH> w;
    int32_t spill_node[wide::kSpill];
    float spill_t[wide::kSpill];
    wide::Stack<LDS_DEPTH> st;
    st.lds = lds + threadIdx.x;
    st.stride = 256;
    st.spill_node = spill_node;
    st.spill_t = spill_t;
    w.run(s, rays, hits, i < n ? i : -1, flags, max_iter, tri_tests, st);
}

template <bool ANY, int W, int LDS_DEPTH, int INNER, int LEAF_BATCH>
__global__ void __launch_bounds__(256) wide_grid(SceneArgs s, const crt_ray* __restrict__ rays,
                                                 crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                                 int32_t max_iter, uint32_t* __restrict__ tri_tests, int32_t blocks) {
    __shared__ int32_t lds[wide::lds_ints<W, LDS_DEPTH>() * 256];
    grid_body<ANY, W, LDS_DEPTH, INNER, LEAF_BATCH>(s, rays, hits, n, flags, max_iter, tri_tests, blocks, lds);
}

// The same kernel compiled for at least OCC waves per SIMD (register budget
// 512 / OCC), trading a few register spills for occupancy.
template <bool ANY, int W, int LDS_DEPTH, int INNER, int LEAF_BATCH, int OCC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC, 8)))
wide_grid_occ(SceneArgs s, const crt_ray* __restrict__ rays, crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
              int32_t max_iter, uint32_t* __restrict__ tri_tests, int32_t blocks) {
    __shared__ int32_t lds[wide::lds_ints<W, LDS_DEPTH>() * 256];
    grid_body<ANY, W, LDS_DEPTH, INNER, LEAF_BATCH>(s, rays, hits, n, flags, max_iter, tri_tests, blocks, lds);
}

// Device-filling waves that claim rays from the sharded queue whenever
// enough of their lanes are idle (wide_walk.h: WideWave::run_refill).
template <bool ANY, int W, int LDS_DEPTH, int INNER, int LEAF_BATCH, int REFILL>
__device__ __forceinline__ void persistent_body(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                                crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                                int32_t max_iter, uint32_t* __restrict__ tri_tests,
                                                unsigned long long* __restrict__ counters, int32_t* lds) {
    int32_t spill_node[wide::kSpill];
    float spill_t[wide::kSpill];
    wide::Stack<LDS_DEPTH> st;
    st.lds = lds + threadIdx.x;
    st.stride = 256;
    st.spill_node = spill_node;
    st.spill_t = spill_t;
    wide::WideWave<ANY, W, LDS_DEPTH, INNER, LEAF_BATCH> w;
    w.run_refill(s, rays, hits, n, flags, max_iter, tri_tests, st, counters, REFILL);
}

template <bool ANY, int W, int LDS_DEPTH, int INNER, int LEAF_BATCH, int REFILL>
__global__ void __launch_bounds__(256) wide_persistent(SceneArgs s, const crt_ray* __restrict__ rays,
                                                       crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                                       int32_t max_iter, uint32_t* __restrict__ tri_tests,
                                                       unsigned long long* __restrict__ counters) {
    __shared__ int32_t lds[wide::lds_ints<W, LDS_DEPTH>() * 256];
    persistent_body<ANY, W, LDS_DEPTH, INNER, LEAF_BATCH, REFILL>(s, rays, hits, n, flags, max_iter, tri_tests,
                                                                  counters, lds);
}

template <bool ANY, int W, int LDS_DEPTH, int INNER, int LEAF_BATCH, int REFILL, int OCC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC, 8)))
wide_persistent_occ(SceneArgs s, const crt_ray* __restrict__ rays, crt_hit* __restrict__ hits, int64_t n,
                    uint32_t flags, int32_t max_iter, uint32_t* __restrict__ tri_tests,
                    unsigned long long* __restrict__ counters) {
    __shared__ int32_t lds[wide::lds_ints<W, LDS_DEPTH>() * 256];
    persistent_body<ANY, W, LDS_DEPTH, INNER, LEAF_BATCH, REFILL>(s, rays, hits, n, flags, max_iter, tri_tests,
                                                                  counters, lds);
}

template <bool ANY, int W, int D, int I, int LB, int OCC>
const void* grid_kernel() {
    if constexpr (OCC == 0) {
        return reinterpret_cast<const void*>(&wide_grid<ANY, W, D, I, LB>);
    } else {
        return reinterpret_cast<const void*>(&wide_grid_occ<ANY, W, D, I, LB, OCC>);
    }
}

template <bool ANY, int W, int D, int I, int LB, int R, int OCC>
const void* persistent_kernel() {
    if constexpr (OCC == 0) {
        return reinterpret_cast<const void*>(&wide_persistent<ANY, W, D, I, LB, R>);
    } else {
        return reinterpret_cast<const void*>(&wide_persistent_occ<ANY, W, D, I, LB, R, OCC>);
    }
}

}  // namespace wdev

// Presets (variant, LDS stack entries, INNER node steps, LEAF_BATCH lanes,
// OCC waves per SIMD the register budget must allow; 0 = the compiler's choice).
#define CRT_WIDE_PRESETS(X) \
    X(0, 8, 4, 16, 0)       \
    X(1, 8, 8, 16, 0)       \
    X(2, 8, 2, 8, 0)        \
    X(3, 16, 4, 16, 0)      \
    X(4, 8, 4, 32, 0)       \
    X(5, 8, 16, 48, 0)      \
    X(6, 8, 4, 32, 6)       \
    X(7, 8, 8, 16, 6)       \
    X(8, 8, 4, 16, 6)       \
    X(9, 8, 16, 48, 6)

template <bool ANY, int W>
static hipError_t launch_wide_w(const SceneArgs& s, const crt_ray* rays, crt_hit* hits, int64_t n, uint32_t flags,
                                int32_t max_iter, uint32_t* tri_tests, int variant, hipStream_t stream) {
    int32_t blocks = int32_t((n + 255) / 256);
    const unsigned padded = unsigned((blocks + 7) / 8 * 8);
    const void* fn = nullptr;
#define CRT_WIDE_CASE(V, D, I, LB, OCC)                      \
    case V:                                                \
        fn = wdev::grid_kernel<ANY, W, D, I, LB, OCC>();    \
        break;
    switch (variant) {
        CRT_WIDE_PRESETS(CRT_WIDE_CASE)
        default: return hipErrorInvalidValue;
    }
#undef CRT_WIDE_CASE
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

// Presets of the refilling kernel (variant, LDS stack entries, INNER, LEAF_BATCH, REFILL idle lanes, OCC).
#define CRT_WIDE_P_PRESETS(X) \
    X(0, 8, 4, 32, 16, 0)     \
    X(1, 8, 4, 16, 8, 0)      \
    X(2, 8, 2, 16, 16, 0)     \
    X(3, 8, 4, 32, 32, 0)     \
    X(4, 8, 8, 32, 16, 0)     \
    X(5, 8, 4, 48, 24, 0)     \
    X(6, 8, 4, 32, 16, 5)     \
    X(7, 8, 4, 16, 8, 5)      \
    X(8, 8, 8, 32, 16, 5)     \
    X(9, 8, 2, 16, 16, 5)

template <bool ANY, int W>
static const void* wide_persistent_fn(int variant) {
#define CRT_WIDE_P_FN(V, D, I, LB, R, OCC) \
    case V: return wdev::persistent_kernel<ANY, W, D, I, LB, R, OCC>();
    switch (variant) {
        CRT_WIDE_P_PRESETS(CRT_WIDE_P_FN)
        default: return nullptr;
    }
#undef CRT_WIDE_P_FN
}

static const void* wide_persistent_kernel(int format, int query, int variant) {
    const bool any = query == CRT_QUERY_ANY;
    if (format == CRT_FORMAT_WIDE8) return any ? wide_persistent_fn<true, 8>(variant) : wide_persistent_fn<false, 8>(variant);
    return any ? wide_persistent_fn<true, 4>(variant) : wide_persistent_fn<false, 4>(variant);
}

int wide_persistent_fill_grid(int format, int query, int variant) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    const void* fn = wide_persistent_kernel(format, query, variant);
    if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess || per_cu <= 0)
        per_cu = 4;
    return cus * per_cu;
}

hipError_t launch_wide_persistent(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                                  uint32_t flags, int32_t max_iter, unsigned long long* counters, uint32_t* tri_tests,
                                  int variant, int grid, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const void* fn = wide_persistent_kernel(s.format, query, variant);
    if (!fn) return hipErrorInvalidValue;
    if (grid <= 0) grid = wide_persistent_fill_grid(s.format, query, variant);
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

hipError_t launch_wide_grid(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                            uint32_t flags, int32_t max_iter, uint32_t* tri_tests, int variant, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const bool any = query == CRT_QUERY_ANY;
    if (s.format == CRT_FORMAT_WIDE8)
        return any ? launch_wide_w<true, 8>(s, rays, hits, n, flags, max_iter, tri_tests, variant, stream)
                   : launch_wide_w<false, 8>(s, rays, hits, n, flags, max_iter, tri_tests, variant, stream);
    return any ? launch_wide_w<true, 4>(s, rays, hits, n, flags, max_iter, tri_tests, variant, stream)
               : launch_wide_w<false, 4>(s, rays, hits, n, flags, max_iter, tri_tests, variant, stream);
}

}  // namespace crt
