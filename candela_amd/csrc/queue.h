// Sharded work queue for the refilling persistent kernels.
//
// A refilling wave claims one index per idle lane with a single atomicAdd.
// With one counter for the whole launch, those atomics serialize on one
// cache line: at about 5.5 ns each, a launch of 1.3M rays refilled 8 lanes at
// a time spends 0.9 ms in the queue alone, more than the walk.  The queue
// below splits [0, n) into kShards contiguous shards with a counter each, on
// its own 128-byte line.  A wave starts on the shard of its XCD (workgroups
// are dealt to the 8 XCDs round robin, two shards per XCD), so its rays are
// neighbours in the input and its atomics go to a line few other waves touch.
// When a shard runs dry the wave moves to the next one; after it has found
// every shard dry it stops claiming.  Every index of every shard is handed
// out exactly once: claims on one counter are disjoint, and no wave leaves a
// shard before a claim on it has passed the shard's end.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace crt {

constexpr int kQueueShards = 16;
constexpr int kQueueStride = 16;   // u64 between counters: 128 bytes
constexpr int kQueueWords = kQueueShards * kQueueStride;

// Zeroes the counters of one launch (one workgroup of kQueueShards lanes).
[[maybe_unused]] static __global__ void reset_queue(unsigned long long* __restrict__ counters) {
    if (threadIdx.x < kQueueShards) counters[threadIdx.x * kQueueStride] = 0ull;
}

struct ShardQueue {
    unsigned long long* counters;
    int64_t n;
    int32_t nsh;   // shards in use, 1..kQueueShards
    int32_t shard, visited, cshard;
    bool strided;   // shard k holds k, k + kQueueShards, ... instead of a contiguous range
    bool more;      // wave-uniform: some shard may still hold indices

    __device__ __forceinline__ void init(unsigned long long* c, int64_t total, bool stride_shards = false,
                                         int32_t shards = kQueueShards) {
        counters = c;
        n = total;
        strided = stride_shards;
        nsh = shards < 1 ? 1 : (shards > kQueueShards ? kQueueShards : shards);
        // wave-uniform values are read into scalar registers (readfirstlane)
        // so that the queue state takes no vector registers
        const int32_t wave = __builtin_amdgcn_readfirstlane(int32_t(threadIdx.x >> 6));
        shard = int32_t((blockIdx.x % 8u) * 2u + uint32_t(wave & 1)) % nsh;
        cshard = shard;
        visited = 0;
        more = total > 0;
    }
    __device__ __forceinline__ int64_t lo(int32_t k) const { return n * k / nsh; }
    __device__ __forceinline__ int64_t size(int32_t k) const {
        if (strided) return n > k ? (n - k + nsh - 1) / nsh : 0;
        return lo(k + 1) - lo(k);
    }
    __device__ __forceinline__ void advance() {
        if (++visited >= nsh) {
            more = false;
        } else {
            shard = shard + 1 == nsh ? 0 : shard + 1;
        }
    }
    // Claims up to `want` positions of one shard for the calling wave (every
    // lane calls with the same `want`).  Returns the first position; `got`
    // (uniform) is how many are valid, 0 once the queue is dry.  at() maps a
    // position of the last claim to its index in [0, n).
    __device__ __forceinline__ int64_t claim(int32_t want, int32_t& got) {
        got = 0;
        while (more) {
            const int64_t len = size(shard);
            unsigned long long base = 0;
            if (__lane_id() == 0) base = atomicAdd(counters + shard * kQueueStride, (unsigned long long)want);
            base = (unsigned long long)__shfl((long long)base, 0);
            const uint32_t b_lo = __builtin_amdgcn_readfirstlane(uint32_t(base));
            const uint32_t b_hi = __builtin_amdgcn_readfirstlane(uint32_t(base >> 32));
            const int64_t first = int64_t((uint64_t(b_hi) << 32) | b_lo);
            if (first < len) {
                const int64_t left = len - first;
                got = left < want ? int32_t(left) : want;
                cshard = shard;
                if (left <= want) advance();
                return first;
            }
            advance();
        }
        return 0;
    }
    __device__ __forceinline__ int64_t at(int64_t pos) const {
        return strided ? int64_t(cshard) + int64_t(nsh) * pos : lo(cshard) + pos;
    }
};

}  // namespace crt
