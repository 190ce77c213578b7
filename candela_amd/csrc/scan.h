// Stable compaction helpers shared by the ray generators (raygen.hip) and the
// device BVH builder (gpu_build.hip): a workgroup ranks its flagged lanes with
// one ballot per wave, one workgroup scans the per-workgroup counts, and each
// workgroup then writes its elements at its offset in input order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace crt {
namespace scan {

constexpr int kBlock = 256;      // workgroup of the count and scatter kernels
constexpr int kScanBlock = 1024; // the single scanning workgroup

// Rank of this lane among the flagged lanes of its workgroup (kBlock
// threads), and the workgroup's count.  wave_counts: LDS of kBlock / 64
// entries.  Every thread of the workgroup must call it.
__device__ __forceinline__ uint32_t block_rank(bool flag, uint32_t* wave_counts, uint32_t& total) {
    const uint64_t ballot = __ballot(flag);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t below =
        __builtin_amdgcn_mbcnt_hi(uint32_t(ballot >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(ballot), 0u));
    if (lane == 0) wave_counts[wave] = uint32_t(__popcll(ballot));
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        const uint32_t c = wave_counts[w];
        before += w < wave ? c : 0u;
        total += c;
    }
    return before + below;
}

// Exclusive scan of nb counts by one workgroup of kScanBlock threads; the
// total goes to *total.
[[maybe_unused]] static __global__ void __launch_bounds__(kScanBlock)
    exclusive_scan(const uint32_t* __restrict__ counts, int64_t nb, int64_t* __restrict__ offsets,
                   int64_t* __restrict__ total) {
    __shared__ int64_t wave_sums[kScanBlock / 64];
    __shared__ int64_t carry_sh;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry_sh = 0;
    __syncthreads();
    for (int64_t base = 0; base < nb; base += kScanBlock) {
        const int64_t k = base + threadIdx.x;
        const int64_t v = k < nb ? int64_t(counts[k]) : 0;
        int64_t incl = v;
        for (int d = 1; d < 64; d <<= 1) {
            const int64_t o = __shfl_up(incl, d, 64);
            if (lane >= d) incl += o;
        }
        if (lane == 63) wave_sums[wave] = incl;
        __syncthreads();
        int64_t before = carry_sh;
        int64_t chunk = 0;
        for (int w = 0; w < kScanBlock / 64; ++w) {
            before += w < wave ? wave_sums[w] : 0;
            chunk += wave_sums[w];
        }
        if (k < nb) offsets[k] = before + incl - v;
        __syncthreads();
        if (threadIdx.x == 0) carry_sh += chunk;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry_sh;
}

}  // namespace scan
}  // namespace crt
