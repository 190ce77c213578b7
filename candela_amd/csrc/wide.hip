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
