// Device BVH builder: a linear BVH over Morton codes (Karras, "Maximizing
// Parallelism in the Construction of BVHs, Octrees, and k-d Trees", HPG
// 2012), written into the reference's stack layout.
//
// The reference builds on the CPU (BuildBVH, BVHConstructor.cpp:951-1108)
// with binned SAH; that tree is not what this builder produces, so its
// buffers are not byte-identical to BuildBVH's (traversal results are: every
// walk reports the nearest hit).  It exists for scenes rebuilt every frame:
// 15.6M triangles take seconds on the host and milliseconds here.
//
// Steps, all on one stream, none allocating (the caller passes a workspace):
//   1. per triangle its box centroid ((min + max) * 0.5, as ConstructTree
//      caches it, BVHConstructor.cpp:404-427); bounds of the centroids by a
//      two-level min/max reduction;
//   2. a 30-bit Morton code of the centroid quantized to 1024 cells per axis;
//   3. a stable LSD radix sort of (code, triangle) by 4-bit digits, 8 passes:
//      each pass counts the digits per 1024-code tile, scans the digit-major
//      counts (a workgroup per digit) and scatters each tile's codes in input
//      order within their digit;
//   4. Karras's radix tree: internal node i finds its range and split by
//      binary searches over the longest common prefix of neighbouring codes
//      (ties broken by the sorted index, so every code is distinct);
//   5. boxes: a node's box is the min / max over the triangles of its leaf
//      range.  Nodes inside one workgroup's 1024 leaves are united bottom up
//      in LDS; the workgroup also writes its leaves' prefix and suffix boxes
//      and lists the nodes whose range spans workgroups.  A wave then forms
//      each listed node's box from one suffix, one prefix and the whole
//      chunks between them (no chain of dependent steps up the tree);
//   6. the stack layout: entry i holds internal node i's two children; a
//      child that covers two leaves is written as one 2-triangle leaf (the
//      reference's leaf size), so its own entry is never reached.
// The triangle array is the input triangles in sorted order.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>

#include "candela_rt.h"
#include "crt_internal.h"
#include "scan.h"

namespace crt {
namespace gb {

using scan::kBlock;
using scan::kScanBlock;

constexpr int kMortonBits = 30;

struct Box {
    float lo[3], hi[3];
};

__device__ __forceinline__ Box tri_box(const crt_vertex* __restrict__ v, const uint32_t* __restrict__ idx, int64_t t) {
    Box b;
    const float4 p0 = *reinterpret_cast<const float4*>(v[idx[3 * t + 0]].position);
    const float4 p1 = *reinterpret_cast<const float4*>(v[idx[3 * t + 1]].position);
    const float4 p2 = *reinterpret_cast<const float4*>(v[idx[3 * t + 2]].position);
    b.lo[0] = fminf(fminf(p0.x, p1.x), p2.x);
    b.lo[1] = fminf(fminf(p0.y, p1.y), p2.y);
    b.lo[2] = fminf(fminf(p0.z, p1.z), p2.z);
    b.hi[0] = fmaxf(fmaxf(p0.x, p1.x), p2.x);
    b.hi[1] = fmaxf(fmaxf(p0.y, p1.y), p2.y);
    b.hi[2] = fmaxf(fmaxf(p0.z, p1.z), p2.z);
    return b;
}

__device__ __forceinline__ void centroid(const Box& b, float c[3]) {
    for (int k = 0; k < 3; ++k) c[k] = (b.lo[k] + b.hi[k]) * 0.5f;
}

// Per-workgroup min / max of the centroids: partial[6 * block + (0..2 lo, 3..5 hi)].
__global__ void __launch_bounds__(kBlock) centroid_bounds(const crt_vertex* __restrict__ v,
                                                          const uint32_t* __restrict__ idx, int64_t n,
                                                          float* __restrict__ partial) {
    __shared__ float sh[6][kBlock];
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    float c[3] = {0.0f, 0.0f, 0.0f};
    const bool ok = t < n;
    if (ok) centroid(tri_box(v, idx, t), c);
    for (int k = 0; k < 3; ++k) {
        sh[k][threadIdx.x] = ok ? c[k] : __builtin_inff();
        sh[3 + k][threadIdx.x] = ok ? c[k] : -__builtin_inff();
    }
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s)
            for (int k = 0; k < 3; ++k) {
                sh[k][threadIdx.x] = fminf(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
                sh[3 + k][threadIdx.x] = fmaxf(sh[3 + k][threadIdx.x], sh[3 + k][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * int64_t(blockIdx.x) + threadIdx.x] = sh[threadIdx.x][0];
}

// One workgroup folds the partials into (lo[3], scale[3]): scale = 1 / extent, 0 for a flat axis.
constexpr int kBoundsBlock = 1024;   // the one workgroup that reduces the partial bounds

__global__ void __launch_bounds__(kBoundsBlock) finish_bounds(const float* __restrict__ partial, int64_t nb,
                                                              float* __restrict__ frame) {
    __shared__ float sh[6][kBoundsBlock];
    float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
#pragma unroll 4
    for (int64_t b = threadIdx.x; b < nb; b += kBoundsBlock)
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], partial[6 * b + k]);
            hi[k] = fmaxf(hi[k], partial[6 * b + 3 + k]);
        }
    for (int k = 0; k < 3; ++k) {
        sh[k][threadIdx.x] = lo[k];
        sh[3 + k][threadIdx.x] = hi[k];
    }
    __syncthreads();
    for (int s = kBoundsBlock / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s)
            for (int k = 0; k < 3; ++k) {
                sh[k][threadIdx.x] = fminf(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
                sh[3 + k][threadIdx.x] = fmaxf(sh[3 + k][threadIdx.x], sh[3 + k][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        const float l = sh[threadIdx.x][0], h = sh[3 + threadIdx.x][0];
        const float ext = h - l;
        frame[threadIdx.x] = l;
        frame[3 + threadIdx.x] = ext > 0.0f ? 1.0f / ext : 0.0f;
    }
}

// 10 bits spread to every third bit.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ uint32_t quantize(float c, float lo, float scale) {
    const float x = ((c - lo) * scale) * 1024.0f;
    return uint32_t(fminf(fmaxf(x, 0.0f), 1023.0f));
}

__global__ void __launch_bounds__(kBlock) morton(const crt_vertex* __restrict__ v, const uint32_t* __restrict__ idx,
                                                 int64_t n, const float* __restrict__ frame,
                                                 uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= n) return;
    float c[3];
    centroid(tri_box(v, idx, t), c);
    const uint32_t x = quantize(c[0], frame[0], frame[3]);
    const uint32_t y = quantize(c[1], frame[1], frame[4]);
    const uint32_t z = quantize(c[2], frame[2], frame[5]);
    keys[t] = (spread3(x) << 2) | (spread3(y) << 1) | spread3(z);
    vals[t] = uint32_t(t);
}

// Radix sort of the codes by 4-bit digits (LSD, stable), tiles of kTile
// codes: a workgroup of kBlock threads takes kItems codes each, round r
// holding codes tile * kTile + r * kBlock + thread, so rounds and threads run
// in index order.  tile_ranks gives each code its rank among the codes of its
// digit that precede it in the tile (peers in the wave from one ballot per
// digit bit; earlier waves and rounds from LDS counts) and leaves the tile's
// per-digit counts in run[].  Codes past n take digit 15: they rank after
// every code of the tile, and only the last tile has them, whose digit-15
// segment ends the order.
constexpr int kDigitBits = 4;
constexpr int kRadix = 1 << kDigitBits;
constexpr int kItems = 4;
constexpr int kTile = kBlock * kItems;
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ void tile_ranks(const uint32_t* __restrict__ keys, int64_t n, int shift, uint32_t key[kItems],
                                           uint32_t digit[kItems], uint32_t rank[kItems], uint32_t* wc,
                                           uint32_t* run) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    if (threadIdx.x < kRadix) run[threadIdx.x] = 0u;
    for (int r = 0; r < kItems; ++r) {
        const int64_t i = int64_t(blockIdx.x) * kTile + r * kBlock + threadIdx.x;
        key[r] = i < n ? keys[i] : 0u;
        const uint32_t d = i < n ? (key[r] >> shift) & (kRadix - 1) : kRadix - 1;
        digit[r] = d;
        uint64_t peers = ~0ull;
        for (int b = 0; b < kDigitBits; ++b) {
            const uint64_t bb = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        if (threadIdx.x < kWaves * kRadix) wc[threadIdx.x] = 0u;
        __syncthreads();
        if ((peers & below) == 0ull) wc[wave * kRadix + d] = uint32_t(__popcll(peers));   // the digit's first lane
        __syncthreads();
        uint32_t before = run[d];
        for (int w = 0; w < kWaves; ++w) before += w < wave ? wc[w * kRadix + d] : 0u;
        rank[r] = before + uint32_t(__popcll(peers & below));
        __syncthreads();
        if (threadIdx.x < kRadix) {
            uint32_t sum = 0u;
            for (int w = 0; w < kWaves; ++w) sum += wc[w * kRadix + threadIdx.x];
            run[threadIdx.x] += sum;
        }
        __syncthreads();
    }
}

// counts[digit * ntiles + tile]: digit-major, so one exclusive scan gives
// each (digit, tile) its first position in the stable order.
__global__ void __launch_bounds__(kBlock) count_digits(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                       uint32_t* __restrict__ counts) {
    __shared__ uint32_t wc[kWaves * kRadix];
    __shared__ uint32_t run[kRadix];
    uint32_t key[kItems], digit[kItems], rank[kItems];
    tile_ranks(keys, n, shift, key, digit, rank, wc, run);
    if (threadIdx.x < kRadix) counts[int64_t(threadIdx.x) * gridDim.x + blockIdx.x] = run[threadIdx.x];
}

// Exclusive scan of the digit-major tile counts (counts[d * nt + tile]) by
// kRadix workgroups, one per digit: workgroup d sums the counts of the digits
// below it (a contiguous prefix of the array) for its base, then scans its
// own nt counts in coalesced passes.  One workgroup scanning all kRadix * nt
// counts ran on one CU and took a third of the build.
constexpr int kDigitScanBlock = 1024;

__global__ void __launch_bounds__(kDigitScanBlock) scan_digits(const uint32_t* __restrict__ counts, int64_t nt,
                                                               int64_t* __restrict__ offsets) {
    __shared__ int64_t wave_sums[kDigitScanBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t seg = int64_t(blockIdx.x) * nt;
    // base: the sum of counts [0, seg)
    int64_t p0 = 0, p1 = 0, p2 = 0, p3 = 0;   // four loads in flight per thread
    for (int64_t k = threadIdx.x; k < seg; k += 4 * kDigitScanBlock) {
        p0 += counts[k];
        p1 += k + kDigitScanBlock < seg ? counts[k + kDigitScanBlock] : 0u;
        p2 += k + 2 * kDigitScanBlock < seg ? counts[k + 2 * kDigitScanBlock] : 0u;
        p3 += k + 3 * kDigitScanBlock < seg ? counts[k + 3 * kDigitScanBlock] : 0u;
    }
    int64_t part = (p0 + p1) + (p2 + p3);
    for (int d = 32; d > 0; d >>= 1) part += __shfl_xor(part, d, 64);
    if (lane == 0) wave_sums[wave] = part;
    __syncthreads();
    int64_t carry = 0;
    for (int w = 0; w < kDigitScanBlock / 64; ++w) carry += wave_sums[w];
    __syncthreads();
    for (int64_t b = 0; b < nt; b += kDigitScanBlock) {
        const int64_t k = b + threadIdx.x;
        const int64_t v = k < nt ? int64_t(counts[seg + k]) : 0;
        int64_t incl = v;
        for (int d = 1; d < 64; d <<= 1) {
            const int64_t o = __shfl_up(incl, d, 64);
            if (lane >= d) incl += o;
        }
        if (lane == 63) wave_sums[wave] = incl;
        __syncthreads();
        int64_t before = carry, chunk = 0;
        for (int w = 0; w < kDigitScanBlock / 64; ++w) {
            const int64_t ws = wave_sums[w];
            before += w < wave ? ws : 0;
            chunk += ws;
        }
        if (k < nt) offsets[seg + k] = before + incl - v;
        carry += chunk;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kBlock) scatter_digits(const uint32_t* __restrict__ kin,
                                                         const uint32_t* __restrict__ vin, int64_t n, int shift,
                                                         const int64_t* __restrict__ offsets,
                                                         uint32_t* __restrict__ kout, uint32_t* __restrict__ vout) {
    __shared__ uint32_t wc[kWaves * kRadix];
    __shared__ uint32_t run[kRadix];
    uint32_t key[kItems], digit[kItems], rank[kItems];
    tile_ranks(kin, n, shift, key, digit, rank, wc, run);
    for (int r = 0; r < kItems; ++r) {
        const int64_t i = int64_t(blockIdx.x) * kTile + r * kBlock + threadIdx.x;
        if (i >= n) continue;
        const int64_t pos = offsets[int64_t(digit[r]) * gridDim.x + blockIdx.x] + rank[r];
        kout[pos] = key[r];
        vout[pos] = vin[i];
    }
}

// Longest common prefix of sorted codes i and j (Karras's delta); -1 outside [0, n).
__device__ __forceinline__ int delta(const uint32_t* __restrict__ keys, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + __clz(uint32_t(i ^ j));
    return __clz(a ^ b);
}

// Internal node i (0 <= i < n - 1): children (c < n - 1: internal c; else leaf c - (n - 1)), covered leaf
// range, and the parent of every node (parent[internal c] at c, parent[leaf l] at n - 1 + l).
__global__ void __launch_bounds__(kBlock) karras(const uint32_t* __restrict__ keys, int64_t n,
                                                 int2* __restrict__ child, int2* __restrict__ range,
                                                 int32_t* __restrict__ parent, unsigned* __restrict__ span_count) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i == 0) *span_count = 0u;   // the boxes pass appends to the span list
    if (i >= n - 1) return;
    const int d = delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int64_t lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int64_t s = 0;
    int64_t t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    const int64_t first = i < j ? i : j, last = i < j ? j : i;
    const int32_t left = int32_t(first == gamma ? (n - 1) + gamma : gamma);
    const int32_t right = int32_t(last == gamma + 1 ? (n - 1) + gamma + 1 : gamma + 1);
    child[i] = make_int2(left, right);
    range[i] = make_int2(int32_t(first), int32_t(last));
    parent[left] = int32_t(i);
    parent[right] = int32_t(i);
}

// Boxes.  A node's box is the min / max over the triangle boxes of its leaf
// range, which min / max give exactly in any order, so the boxes equal a
// bottom-up union's bit for bit.
//
// Karras's tree gives internal node i a leaf range with one end at i, so a
// node whose range lies in one workgroup's kBoxBlock leaves has its index in
// that workgroup's leaf index range, and its whole subtree is computed by that
// workgroup: a thread per leaf climbs while it is the second of a node's two
// children to arrive (counters and boxes in LDS, workgroup scope).  The same
// pass writes each leaf's prefix box (the chunk's leaves up to it) and suffix
// box (from it to the chunk's end) and lists, a thread per internal node, the
// nodes whose range spans chunks.  The span pass then gives each listed node
// [a, b] the union of suffix(a), prefix(b) and the whole chunks between them
// (a chunk's box is the suffix box of its first leaf).  No pass has a chain
// of dependent steps through the upper tree, and none needs agent-scope
// fences, which on this chip write back and invalidate an XCD's L2.
constexpr int kBoxBlock = 1024;
constexpr int kSpanGrid = 512;   // workgroups of the span pass (kBlock threads, a node per wave)

__device__ __forceinline__ Box unite(const Box& a, const Box& b) {
    Box u;
    for (int q = 0; q < 3; ++q) {
        u.lo[q] = fminf(a.lo[q], b.lo[q]);
        u.hi[q] = fmaxf(a.hi[q], b.hi[q]);
    }
    return u;
}

__device__ __forceinline__ Box empty_box() {
    Box b;
    for (int q = 0; q < 3; ++q) {
        b.lo[q] = __builtin_inff();
        b.hi[q] = -__builtin_inff();
    }
    return b;
}

__device__ __forceinline__ Box shfl_box(const Box& b, int src_lane) {
    Box o;
    for (int q = 0; q < 3; ++q) {
        o.lo[q] = __shfl(b.lo[q], src_lane, 64);
        o.hi[q] = __shfl(b.hi[q], src_lane, 64);
    }
    return o;
}

__global__ void __launch_bounds__(kBoxBlock) boxes_local(const crt_vertex* __restrict__ v,
                                                         const uint32_t* __restrict__ idx,
                                                         const uint32_t* __restrict__ order, int64_t n,
                                                         const int2* __restrict__ child, const int2* __restrict__ range,
                                                         const int32_t* __restrict__ parent, Box* __restrict__ boxes,
                                                         Box* __restrict__ pre, Box* __restrict__ suf,
                                                         Box* __restrict__ leaf_box, int32_t* __restrict__ span,
                                                         unsigned* __restrict__ span_count) {
    __shared__ unsigned lflag[kBoxBlock];
    __shared__ Box lbox[kBoxBlock];
    __shared__ Box ltri[kBoxBlock];   // the chunk's leaf boxes, for the climb
    __shared__ Box wave_pre[kBoxBlock / 64], wave_suf[kBoxBlock / 64];
    const int64_t base = int64_t(blockIdx.x) * kBoxBlock;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t l = base + threadIdx.x;
    const bool valid = l < n;
    lflag[threadIdx.x] = 0u;
    // prefix and suffix boxes of the chunk: wave scans, then the wave totals
    const Box tb = valid ? tri_box(v, idx, int64_t(order[l])) : empty_box();
    ltri[threadIdx.x] = tb;
    Box p = tb, q = tb;
    for (int d = 1; d < 64; d <<= 1) {
        const Box op = shfl_box(p, lane >= d ? lane - d : lane);
        const Box oq = shfl_box(q, lane + d < 64 ? lane + d : lane);
        if (lane >= d) p = unite(p, op);
        if (lane + d < 64) q = unite(q, oq);
    }
    if (lane == 63) wave_pre[wave] = p;
    if (lane == 0) wave_suf[wave] = q;
    __syncthreads();
    for (int w = 0; w < kBoxBlock / 64; ++w) {
        if (w < wave) p = unite(p, wave_pre[w]);
        if (w > wave) q = unite(q, wave_suf[w]);
    }
    // internal node l (l < n - 1) goes to the span list when its range leaves
    // one chunk: a ballot per wave, one append per wave
    bool spans = false;
    if (l < n - 1) {
        const int2 r = range[l];
        spans = r.x / kBoxBlock != r.y / kBoxBlock;
    }
    const uint64_t ballot = __ballot(spans);
    if (ballot) {
        unsigned at = 0;
        if (lane == 0)
            at = __hip_atomic_fetch_add(span_count, unsigned(__popcll(ballot)), __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
        at = __shfl(at, 0, 64);
        const unsigned below =
            __builtin_amdgcn_mbcnt_hi(uint32_t(ballot >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(ballot), 0u));
        if (spans) span[at + below] = int32_t(l);
    }
    if (!valid) return;   // no barrier below
    pre[l] = p;
    suf[l] = q;
    leaf_box[l] = tb;   // emit writes the leaf children from these
    int32_t node = parent[(n - 1) + l];
    while (node >= 0) {
        const int2 r = range[node];
        if (r.x / kBoxBlock != r.y / kBoxBlock) return;   // the span pass forms it
        const int32_t li = node - int32_t(base);   // in [0, kBoxBlock): node lies in this workgroup's leaf range
        if (__hip_atomic_fetch_add(&lflag[li], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u) return;
        const int2 c = child[node];
        // a local node's children lie in this chunk: leaves in ltri, nodes in lbox
        const Box b0 = c.x >= n - 1 ? ltri[c.x - int32_t(n - 1) - int32_t(base)] : lbox[c.x - int32_t(base)];
        const Box b1 = c.y >= n - 1 ? ltri[c.y - int32_t(n - 1) - int32_t(base)] : lbox[c.y - int32_t(base)];
        const Box u = unite(b0, b1);
        lbox[li] = u;
        boxes[node] = u;
        node = node == 0 ? -1 : parent[node];
    }
}

// A wave per listed node: suffix(a), prefix(b) and the whole chunks of
// [a, b] between them, united across the wave.
__global__ void __launch_bounds__(kBlock) boxes_span(const int2* __restrict__ range, const int32_t* __restrict__ span,
                                                     const unsigned* __restrict__ span_count,
                                                     const Box* __restrict__ pre, const Box* __restrict__ suf,
                                                     Box* __restrict__ boxes) {
    const unsigned count = *span_count;
    const int lane = threadIdx.x & 63;
    const unsigned waves = gridDim.x * (kBlock / 64);
    for (unsigned e = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); e < count; e += waves) {
        const int32_t node = span[e];
        const int2 r = range[node];
        const int64_t ca = r.x / kBoxBlock, cb = r.y / kBoxBlock;
        Box acc = lane == 0 ? suf[r.x] : lane == 1 ? pre[r.y] : empty_box();
        for (int64_t k = ca + 1 + lane; k < cb; k += 64) acc = unite(acc, suf[k * kBoxBlock]);
        for (int d = 32; d > 0; d >>= 1) acc = unite(acc, shfl_box(acc, lane ^ d));
        if (lane == 0) boxes[node] = acc;
    }
}

// Packed leaf word (start << 4 | count) as the stack layout stores it.
__device__ __forceinline__ int32_t leaf_word(int64_t start, int32_t count) {
    return int32_t((uint32_t(start) << 4) | uint32_t(count));
}

__device__ __forceinline__ void put_child(crt_bounds& out, const Box& b, int32_t w_min, int32_t w_max) {
    for (int q = 0; q < 3; ++q) {
        out.min[q] = b.lo[q];
        out.max[q] = b.hi[q];
    }
    out.min[3] = __int_as_float(w_min);
    out.max[3] = __int_as_float(w_max);
}

// Entry i of the stack layout from internal node i; the sorted triangle array.
__global__ void __launch_bounds__(kBlock) emit(const crt_vertex* __restrict__ v, const uint32_t* __restrict__ idx,
                                               const int32_t* __restrict__ mesh, int32_t mesh_all,
                                               const uint32_t* __restrict__ order, int64_t n, int32_t t_offset,
                                               const int2* __restrict__ child, const int2* __restrict__ range,
                                               const Box* __restrict__ boxes, const Box* __restrict__ leaf_box,
                                               crt_stack_node* __restrict__ nodes, crt_triangle* __restrict__ tris) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = order[i];
    crt_triangle tr;
    tr.packed[0] = int32_t(idx[3 * int64_t(t) + 0]);
    tr.packed[1] = int32_t(idx[3 * int64_t(t) + 1]);
    tr.packed[2] = int32_t(idx[3 * int64_t(t) + 2]);
    tr.packed[3] = mesh ? mesh[t] : mesh_all;
    *reinterpret_cast<int4*>(tris + i) = *reinterpret_cast<const int4*>(tr.packed);
    if (i >= n - 1) return;
    const int2 c = child[i];
    crt_stack_node e;
    for (int k = 0; k < 2; ++k) {
        const int32_t cc = k == 0 ? c.x : c.y;
        crt_bounds& out = k == 0 ? e.left : e.right;
        if (cc >= n - 1) {
            const int32_t leaf = cc - int32_t(n - 1);
            put_child(out, leaf_box[leaf], leaf_word(int64_t(t_offset) + leaf, 1), 0);
        } else if (range[cc].y - range[cc].x == 1) {
            put_child(out, boxes[cc], leaf_word(int64_t(t_offset) + range[cc].x, 2), 0);
        } else {
            put_child(out, boxes[cc], -1, cc);
        }
    }
    nodes[i] = e;
}

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct Layout {
    size_t keys[2], vals[2], counts, offsets, partial, frame, child, range, parent, boxes, pre, suf, leaf_box,
        span, span_count, bytes;
};

inline Layout layout(int64_t n) {
    const size_t nb = size_t((n + kBlock - 1) / kBlock);
    const size_t un = size_t(n);
    Layout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = align256(off + bytes);
        return at;
    };
    L.keys[0] = take(un * 4);
    L.keys[1] = take(un * 4);
    L.vals[0] = take(un * 4);
    L.vals[1] = take(un * 4);
    const size_t nt = size_t((n + kTile - 1) / kTile);
    L.counts = take(kRadix * nt * 4);
    L.offsets = take(kRadix * nt * 8);
    L.partial = take(nb * 6 * 4);
    L.frame = take(6 * 4);
    L.child = take(un * 8);
    L.range = take(un * 8);
    L.parent = take(2 * un * 4);
    L.boxes = take(un * sizeof(Box));
    L.pre = take(un * sizeof(Box));
    L.suf = take(un * sizeof(Box));
    L.leaf_box = take(un * sizeof(Box));
    L.span = take(un * 4);
    L.span_count = take(4);
    L.bytes = off;
    return L;
}

}  // namespace gb
}  // namespace crt

extern "C" {

size_t crt_build_bvh_gpu_workspace(int64_t ntris) {
    if (ntris < 2) return 0;
    return crt::gb::layout(ntris).bytes;
}

int crt_build_bvh_gpu(const crt_vertex* vertices, int64_t nverts, const uint32_t* indices, int64_t ntris,
                      const int32_t* mesh_ids, int32_t mesh_id, int32_t t_offset, crt_stack_node* nodes,
                      crt_triangle* triangles, void* workspace, size_t workspace_bytes, void* stream) {
    using namespace crt::gb;
    if (!vertices || !indices || !nodes || !triangles || !workspace)
        return crt::set_error(CRT_E_INVALID, "crt_build_bvh_gpu: null argument");
    if (ntris < 2) return crt::set_error(CRT_E_INVALID, "crt_build_bvh_gpu: the stack layout needs >= 2 triangles");
    if (nverts <= 0 || nverts > 0xFFFFFFFFll) return crt::set_error(CRT_E_INVALID, "crt_build_bvh_gpu: bad vertex count");
    if (t_offset < 0 || uint64_t(ntris) + uint64_t(t_offset) >= (1ull << 28))
        return crt::set_error(CRT_E_INVALID, "crt_build_bvh_gpu: t_offset + triangles exceeds the 28-bit leaf start");
    const Layout L = layout(ntris);
    if (workspace_bytes < L.bytes || reinterpret_cast<uintptr_t>(workspace) % 256 != 0)
        return crt::set_error(CRT_E_INVALID, "crt_build_bvh_gpu: workspace must be 256-byte aligned and hold "
                                             "crt_build_bvh_gpu_workspace(ntris) = " + std::to_string(L.bytes) +
                                             " bytes");
    char* ws = static_cast<char*>(workspace);
    auto at = [&](size_t off) { return static_cast<void*>(ws + off); };
    uint32_t* keys[2] = {static_cast<uint32_t*>(at(L.keys[0])), static_cast<uint32_t*>(at(L.keys[1]))};
    uint32_t* vals[2] = {static_cast<uint32_t*>(at(L.vals[0])), static_cast<uint32_t*>(at(L.vals[1]))};
    auto* counts = static_cast<uint32_t*>(at(L.counts));
    auto* offsets = static_cast<int64_t*>(at(L.offsets));
    auto* partial = static_cast<float*>(at(L.partial));
    auto* frame = static_cast<float*>(at(L.frame));
    auto* child = static_cast<int2*>(at(L.child));
    auto* range = static_cast<int2*>(at(L.range));
    auto* parent = static_cast<int32_t*>(at(L.parent));
    auto* boxes = static_cast<Box*>(at(L.boxes));
    auto* pre = static_cast<Box*>(at(L.pre));
    auto* suf = static_cast<Box*>(at(L.suf));
    auto* leaf_box = static_cast<Box*>(at(L.leaf_box));
    auto* span = static_cast<int32_t*>(at(L.span));
    auto* span_count = static_cast<unsigned*>(at(L.span_count));
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t n = ntris;
    const unsigned nb = unsigned((n + kBlock - 1) / kBlock);

    hipLaunchKernelGGL(centroid_bounds, dim3(nb), dim3(kBlock), 0, st, vertices, indices, n, partial);
    hipLaunchKernelGGL(finish_bounds, dim3(1), dim3(kBoundsBlock), 0, st, partial, int64_t(nb), frame);
    hipLaunchKernelGGL(morton, dim3(nb), dim3(kBlock), 0, st, vertices, indices, n, frame, keys[0], vals[0]);
    int cur = 0;
    const unsigned nt = unsigned((n + kTile - 1) / kTile);
    for (int shift = 0; shift < kMortonBits; shift += kDigitBits) {
        hipLaunchKernelGGL(count_digits, dim3(nt), dim3(kBlock), 0, st, keys[cur], n, shift, counts);
        hipLaunchKernelGGL(scan_digits, dim3(kRadix), dim3(kDigitScanBlock), 0, st, counts, int64_t(nt), offsets);
        hipLaunchKernelGGL(scatter_digits, dim3(nt), dim3(kBlock), 0, st, keys[cur], vals[cur], n, shift, offsets,
                           keys[1 - cur], vals[1 - cur]);
        cur = 1 - cur;
    }
    const unsigned ni = unsigned((n - 1 + kBlock - 1) / kBlock);
    // karras also zeroes the span count (a kernel store, no memset node, so
    // the chain stays capturable)
    hipLaunchKernelGGL(karras, dim3(ni), dim3(kBlock), 0, st, keys[cur], n, child, range, parent, span_count);
    hipLaunchKernelGGL(boxes_local, dim3(unsigned((n + kBoxBlock - 1) / kBoxBlock)), dim3(kBoxBlock), 0, st, vertices,
                       indices, vals[cur], n, child, range, parent, boxes, pre, suf, leaf_box, span, span_count);
    hipLaunchKernelGGL(boxes_span, dim3(kSpanGrid), dim3(kBlock), 0, st, range, span, span_count, pre, suf, boxes);
    hipLaunchKernelGGL(emit, dim3(nb), dim3(kBlock), 0, st, vertices, indices, mesh_ids, mesh_id, vals[cur], n,
                       t_offset, child, range, boxes, leaf_box, nodes, triangles);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return crt::set_error(CRT_E_HIP, std::string("crt_build_bvh_gpu: kernel launch: ") + hipGetErrorString(e));
    return CRT_OK;
}

}  // extern "C"
