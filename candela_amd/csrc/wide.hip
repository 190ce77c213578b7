// Wave-scheduled walk of the wide layout (BVH4 / BVH8): one ray per lane,
// leaves parked and tested by the wave together (wide_walk.h: WideWave), and
// the XCD-aware block order of the stackless grid kernel: blocks that trace
// consecutive ray tiles run on one XCD, whose L2 then holds the part of the
// tree those rays share.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"
#include "trace_launch.h"
#include "wide_walk.h"

#pragma clang fp contract(off)

namespace crt {
namespace wdev {

template <bool ANY, int W, int LDS_DEPTH, int INNER, int LEAF_BATCH>
__global__ void __launch_bounds__(256) wide_grid(SceneArgs s, const crt_ray* __restrict__ rays,
                                                 crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                                 int32_t max_iter, uint32_t* __restrict__ tri_tests, int32_t blocks) {
    __shared__ int32_t lds[wide::lds_ints<W, LDS_DEPTH>() * 256];
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

// Device-filling waves that claim rays from a counter whenever enough of
// their lanes are idle (wide_walk.h: WideWave::run_refill).
template <bool ANY, int W, int LDS_DEPTH, int INNER, int LEAF_BATCH, int REFILL>
__global__ void __launch_bounds__(256) wide_persistent(SceneArgs s, const crt_ray* __restrict__ rays,
                                                       crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                                       int32_t max_iter, uint32_t* __restrict__ tri_tests,
                                                       unsigned long long* __restrict__ counter) {
    __shared__ int32_t lds[wide::lds_ints<W, LDS_DEPTH>() * 256];
    int32_t spill_node[wide::kSpill];
    float spill_t[wide::kSpill];
    wide::Stack<LDS_DEPTH> st;
    st.lds = lds + threadIdx.x;
    st.stride = 256;
    st.spill_node = spill_node;
    st.spill_t = spill_t;
    wide::WideWave<ANY, W, LDS_DEPTH, INNER, LEAF_BATCH> w;
    w.run_refill(s, rays, hits, n, flags, max_iter, tri_tests, st, counter, REFILL);
}

__global__ void reset_counter64(unsigned long long* c) { *c = 0ull; }

}  // namespace wdev

// Presets (variant, LDS stack entries, INNER node steps, LEAF_BATCH lanes).
#define CRT_WIDE_PRESETS(X) \
    X(0, 8, 4, 16)          \
    X(1, 8, 8, 16)          \
    X(2, 8, 2, 8)           \
    X(3, 16, 4, 16)         \
    X(4, 8, 4, 32)          \
    X(5, 8, 16, 48)

template <bool ANY, int W>
static hipError_t launch_wide_w(const SceneArgs& s, const crt_ray* rays, crt_hit* hits, int64_t n, uint32_t flags,
                                int32_t max_iter, uint32_t* tri_tests, int variant, hipStream_t stream) {
    const int32_t blocks = int32_t((n + 255) / 256);
    const unsigned padded = unsigned((blocks + 7) / 8 * 8);
#define CRT_WIDE_CASE(V, D, I, LB)                                                                              \
    case V:                                                                                                   \
        hipLaunchKernelGGL((wdev::wide_grid<ANY, W, D, I, LB>), dim3(padded), dim3(256), 0, stream, s, rays,  \
                           hits, n, flags, max_iter, tri_tests, blocks);                                      \
        break;
    switch (variant) {
        CRT_WIDE_PRESETS(CRT_WIDE_CASE)
        default: return hipErrorInvalidValue;
    }
#undef CRT_WIDE_CASE
    return hipGetLastError();
}

// Presets of the refilling kernel (variant, LDS stack entries, INNER, LEAF_BATCH, REFILL idle lanes).
#define CRT_WIDE_P_PRESETS(X) \
    X(0, 8, 4, 32, 16)        \
    X(1, 8, 4, 16, 8)         \
    X(2, 8, 2, 16, 16)        \
    X(3, 8, 4, 32, 32)        \
    X(4, 8, 8, 32, 16)        \
    X(5, 8, 4, 48, 24)

template <bool ANY, int W>
static const void* wide_persistent_fn(int variant) {
#define CRT_WIDE_P_FN(V, D, I, LB, R) \
    case V: return reinterpret_cast<const void*>(&wdev::wide_persistent<ANY, W, D, I, LB, R>);
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
                                  uint32_t flags, int32_t max_iter, unsigned long long* counter, uint32_t* tri_tests,
                                  int variant, int grid, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const void* fn = wide_persistent_kernel(s.format, query, variant);
    if (!fn) return hipErrorInvalidValue;
    if (grid <= 0) grid = wide_persistent_fill_grid(s.format, query, variant);
    const int64_t need = (n + 255) / 256;
    if (need < grid) grid = int(need);
    hipLaunchKernelGGL(wdev::reset_counter64, dim3(1), dim3(1), 0, stream, counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    SceneArgs sa = s;
    const crt_ray* r = rays;
    crt_hit* h = hits;
    int64_t nn = n;
    uint32_t fl = flags;
    int32_t mi = max_iter;
    uint32_t* tt = tri_tests;
    unsigned long long* c = counter;
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
