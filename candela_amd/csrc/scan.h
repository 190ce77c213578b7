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
// total goes to *total.  A pass covers kScanBlock * kScanPer counts: each
// thread scans kScanPer consecutive ones in registers, the workgroup scans
// the per-thread sums, and the running carry links the passes.
constexpr int kScanPer = 16;

[[maybe_unused]] static __global__ void __launch_bounds__(kScanBlock)
    exclusive_scan(const uint32_t* __restrict__ counts, int64_t nb, int64_t* __restrict__ offsets,
                   int64_t* __restrict__ total) {
    __shared__ int64_t wave_sums[kScanBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t carry = 0;
    for (int64_t pass = 0; pass < nb; pass += int64_t(kScanBlock) * kScanPer) {
        const int64_t base = pass + int64_t(threadIdx.x) * kScanPer;
        int64_t loc[kScanPer];
        int64_t run = 0;
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) {
            loc[j] = run;
            run += base + j < nb ? int64_t(counts[base + j]) : 0;
        }
        int64_t incl = run;
        for (int d = 1; d < 64; d <<= 1) {
            const int64_t o = __shfl_up(incl, d, 64);
            if (lane >= d) incl += o;
        }
        if (lane == 63) wave_sums[wave] = incl;
        __syncthreads();
        int64_t before = carry, chunk = 0;
        for (int w = 0; w < kScanBlock / 64; ++w) {
            const int64_t ws = wave_sums[w];
            before += w < wave ? ws : 0;
            chunk += ws;
        }
        before += incl - run;
#pragma unroll
        for (int j = 0; j < kScanPer; ++j)
            if (base + j < nb) offsets[base + j] = before + loc[j];
        carry += chunk;
        __syncthreads();   // wave_sums is rewritten by the next pass
    }
    if (threadIdx.x == 0) *total = carry;
}

}  // namespace scan
}  // namespace crt
