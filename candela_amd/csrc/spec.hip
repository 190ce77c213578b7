// Launchers of the speculative wave walk of the wide layouts (spec_walk.h):
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
#include "spec_walk.h"
#include "trace_launch.h"

#pragma clang fp contract(off)

namespace crt {
namespace sp {

// Queue slots per lane: the preset's Q for BVH4, Q + 4 (at most 16) for BVH8,
// so that a lane keeps stepping after a node with leaves.
template <int W, int Q>
constexpr int queue_slots() {
    return W == 4 ? Q : (Q + 4 < 16 ? Q + 4 : 16);
}

template <int W, int D, int Q>
struct GroupLds {
    int32_t cols[spec::lds_ints<D, queue_slots<W, Q>()>() * 256];
    spec::WaveLds<queue_slots<W, Q>()> waves[4];
};

template <bool ANY, int W, int D, int I, int Q>
__device__ __forceinline__ void grid_body(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                          crt_hit* __restrict__ hits, int64_t n, uint32_t flags, int32_t max_iter,
                                          uint32_t* __restrict__ tri_tests, int32_t blocks,
                                          GroupLds<W, D, Q>& lds) {
    const int32_t per = (blocks + 7) / 8;  // the grid is padded to a multiple of 8
    const int32_t logical = int32_t(blockIdx.x % 8u) * per + int32_t(blockIdx.x / 8u);
    if (logical >= blocks) return;
    const int64_t i = int64_t(logical) * 256 + threadIdx.x;
    int32_t spill_node[wide::kSpill];
    float spill_t[wide::kSpill];
    wide::Stack<D> st;
    st.lds = lds.cols + threadIdx.x;
    st.stride = 256;
    st.spill_node = spill_node;
    st.spill_t = spill_t;
    spec::SpecWave<ANY, W, D, I, queue_slots<W, Q>()> w;
    w.run(s, rays, hits, i < n ? i : -1, flags, max_iter, tri_tests, st, lds.waves[threadIdx.x >> 6]);
}

template <bool ANY, int W, int D, int I, int Q, int OCC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1, 8)))
spec_grid(SceneArgs s, const crt_ray* __restrict__ rays, crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
          int32_t max_iter, uint32_t* __restrict__ tri_tests, int32_t blocks) {
    __shared__ GroupLds<W, D, Q> lds;
    grid_body<ANY, W, D, I, Q>(s, rays, hits, n, flags, max_iter, tri_tests, blocks, lds);
}

template <bool ANY, int W, int D, int I, int Q, int REFILL, int OCC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1, 8)))
spec_persistent(SceneArgs s, const crt_ray* __restrict__ rays, crt_hit* __restrict__ hits, int64_t n,
                uint32_t flags, int32_t max_iter, uint32_t* __restrict__ tri_tests,
                unsigned long long* __restrict__ counters) {
    __shared__ GroupLds<W, D, Q> lds;
    int32_t spill_node[wide::kSpill];
    float spill_t[wide::kSpill];
    wide::Stack<D> st;
    st.lds = lds.cols + threadIdx.x;
    st.stride = 256;
    st.spill_node = spill_node;
    st.spill_t = spill_t;
    spec::SpecWave<ANY, W, D, I, queue_slots<W, Q>()> w;
    w.run_refill(s, rays, hits, n, flags, max_iter, tri_tests, st, lds.waves[threadIdx.x >> 6], counters, REFILL);
}

}  // namespace sp

// Presets of the grid walk (variant, LDS stack entries, INNER node steps per
// round, Q queue slots per lane (BVH4), OCC waves per SIMD).  Preset 5 keeps
// the workgroup's LDS under 26 KiB so that six workgroups share a CU.
#define CRT_SPEC_PRESETS(X) \
    X(0, 8, 4, 8, 0)        \
    X(1, 8, 8, 8, 0)        \
    X(2, 8, 2, 8, 5)        \
    X(3, 8, 4, 8, 5)        \
    X(4, 8, 8, 8, 5)        \
    X(5, 7, 6, 6, 6)

// Presets of the refilling walk (variant, LDS stack entries, INNER, Q,
// REFILL idle lanes, OCC).
#define CRT_SPEC_P_PRESETS(X)   \
    X(0, 8, 4, 8, 16, 0)        \
    X(1, 8, 8, 8, 16, 0)        \
    X(2, 8, 8, 8, 8, 5)         \
    X(3, 8, 6, 8, 16, 5)        \
    X(4, 8, 8, 8, 16, 5)        \
    X(5, 7, 8, 6, 16, 6)

template <bool ANY, int W>
static const void* spec_grid_fn(int variant) {
#define CRT_SPEC_G_FN(V, D, I, Q, OCC) \
    case V: return reinterpret_cast<const void*>(&sp::spec_grid<ANY, W, D, I, Q, OCC>);
    switch (variant) {
        CRT_SPEC_PRESETS(CRT_SPEC_G_FN)
        default: return nullptr;
    }
#undef CRT_SPEC_G_FN
}

template <bool ANY, int W>
static const void* spec_persistent_fn(int variant) {
#define CRT_SPEC_P_FN(V, D, I, Q, R, OCC) \
    case V: return reinterpret_cast<const void*>(&sp::spec_persistent<ANY, W, D, I, Q, R, OCC>);
    switch (variant) {
        CRT_SPEC_P_PRESETS(CRT_SPEC_P_FN)
        default: return nullptr;
    }
#undef CRT_SPEC_P_FN
}

static const void* pick(bool persistent, int format, int query, int variant) {
    const bool any = query == CRT_QUERY_ANY;
    if (format == CRT_FORMAT_WIDE8)
        return persistent ? (any ? spec_persistent_fn<true, 8>(variant) : spec_persistent_fn<false, 8>(variant))
                          : (any ? spec_grid_fn<true, 8>(variant) : spec_grid_fn<false, 8>(variant));
    return persistent ? (any ? spec_persistent_fn<true, 4>(variant) : spec_persistent_fn<false, 4>(variant))
                      : (any ? spec_grid_fn<true, 4>(variant) : spec_grid_fn<false, 4>(variant));
}

hipError_t launch_spec_grid(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                            uint32_t flags, int32_t max_iter, uint32_t* tri_tests, int variant, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const void* fn = pick(false, s.format, query, variant);
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

hipError_t launch_spec_persistent(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                                  uint32_t flags, int32_t max_iter, unsigned long long* counters,
                                  uint32_t* tri_tests, int variant, int grid, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const void* fn = pick(true, s.format, query, variant);
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
