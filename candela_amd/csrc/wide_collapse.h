// Per-node logic of the device BVH4 / BVH8 collapse (wide_gpu.hip): one
// wide node's slots from its BVH2 stack entry, and its children's numbers.
// Kept free of HIP headers so the same functions also build on the host
// (the collapse tests compile them with clang++ under the undefined-behaviour
// sanitizer, with every local initialized to a pattern, and compare the
// layout with crt_wide_build's); the including file provides __device__,
// __forceinline__, atomicOr, __float_as_int, min, max, fminf and fmaxf.
#pragma once

#include <cstdint>

#include "candela_rt.h"

namespace crt {
namespace wg {

// A BVH2 child as its parent's stack entry stores it, with its leaf range
// when it is (or collapses into) a leaf.
struct Kid {
    float lo[3], hi[3];
    int32_t entry;          // its own stack entry when inner, else -1
    int32_t start, count;   // leaf range when entry < 0
};

// Bits of *err: a link leaves the node array; the links reach more wide
// nodes than the output holds (they do not form a tree); a vertex index
// leaves the vertex array; a leaf range is empty or leaves [0, ntris) (a tree
// built with a triangle offset, or damaged entries).
constexpr unsigned kErrLink = 1u, kErrCap = 2u, kErrVertex = 4u, kErrLeaf = 16u;

// A leaf's triangle range indexes the records: it must lie in [0, ntris).
__device__ __forceinline__ bool leaf_ok(int32_t start, int32_t count, int64_t ntris) {
    return count >= 1 && start >= 0 && int64_t(start) + count <= ntris;
}

// A child slot that fails a check (bad link, bad leaf range) comes back as a
// leaf of count 0, which expand_one writes as an empty slot (count -1).
__device__ __forceinline__ Kid make_kid(const crt_stack_node* __restrict__ nodes, int64_t node_count, int64_t ntris,
                                        const crt_bounds& b, unsigned* err) {
    Kid k;
    for (int q = 0; q < 3; ++q) {
        k.lo[q] = b.min[q];
        k.hi[q] = b.max[q];
    }
    const int32_t wmin = __float_as_int(b.min[3]);
    k.entry = -1;
    k.start = 0;
    k.count = 0;
    if (wmin != -1) {
        k.start = wmin >> 4;
        k.count = wmin & 0xF;
        if (!leaf_ok(k.start, k.count, ntris)) {
            atomicOr(err, kErrLeaf);
            k.start = k.count = 0;
        }
        return k;
    }
    const int32_t e = __float_as_int(b.max[3]);
    if (e < 0 || int64_t(e) >= node_count) {
        atomicOr(err, kErrLink);
        return k;   // read as an empty leaf; the call fails
    }
    // an inner node of two leaves with two adjacent triangles in all is one
    // 2-triangle leaf (the host's rule: at most 2 triangles, contiguous)
    const crt_stack_node& s = nodes[e];
    const int32_t a = __float_as_int(s.left.min[3]), c = __float_as_int(s.right.min[3]);
    if (a != -1 && c != -1) {
        const int32_t sa = a >> 4, na = a & 0xF, sc = c >> 4, nc = c & 0xF;
        const int32_t first = min(sa, sc), last = max(sa + na, sc + nc);
        if (na + nc <= 2 && last - first == na + nc) {
            k.start = first;
            k.count = na + nc;
            if (!leaf_ok(k.start, k.count, ntris)) {
                atomicOr(err, kErrLeaf);
                k.start = k.count = 0;
            }
            return k;
        }
    }
    k.entry = e;
    return k;
}

// The host's half_area: dx * dy + dy * dz + dz * dx, left to right.
__device__ __forceinline__ float half_area(const Kid& k) {
    const float dx = k.hi[0] - k.lo[0], dy = k.hi[1] - k.lo[1], dz = k.hi[2] - k.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

// The slots of one wide node w, the expansion of stack entry `entry`: writes
// its groups and kids[slot] (the stack entry of an inner child, -1
// otherwise); returns its number of inner children.
template <int W>
__device__ __forceinline__ uint32_t expand_one(const crt_stack_node* __restrict__ nodes, int64_t node_count,
                                               int64_t ntris, int32_t entry, int64_t w, crt_wide_group* __restrict__ groups,
                                               int32_t* __restrict__ kids_out, unsigned* __restrict__ err) {
    const crt_stack_node& e = nodes[entry];
    Kid kids[W];
    int n = 0;
    kids[n++] = make_kid(nodes, node_count, ntris, e.left, err);
    kids[n++] = make_kid(nodes, node_count, ntris, e.right, err);
    while (n < W) {
        int best = -1;
        float best_area = -1.0f;
        for (int k = 0; k < n; ++k) {
            if (kids[k].entry < 0) continue;
            const float a = half_area(kids[k]);
            if (a > best_area) {
                best_area = a;
                best = k;
            }
        }
        if (best < 0) break;
        const crt_stack_node& c = nodes[kids[best].entry];
        for (int k = n; k > best + 1; --k) kids[k] = kids[k - 1];
        kids[best + 1] = make_kid(nodes, node_count, ntris, c.right, err);
        kids[best] = make_kid(nodes, node_count, ntris, c.left, err);
        ++n;
    }
    uint32_t m = 0;
    for (int s = 0; s < W; ++s) {
        crt_wide_group& g = groups[w * (W / 4) + s / 4];
        const int j = s % 4;
        const bool used = s < n;
        const Kid& k = kids[used ? s : 0];
        g.lo_x[j] = used ? k.lo[0] : 0.0f;
        g.lo_y[j] = used ? k.lo[1] : 0.0f;
        g.lo_z[j] = used ? k.lo[2] : 0.0f;
        g.hi_x[j] = used ? k.hi[0] : 0.0f;
        g.hi_y[j] = used ? k.hi[1] : 0.0f;
        g.hi_z[j] = used ? k.hi[2] : 0.0f;
        const bool inner = used && k.entry >= 0;
        const bool leaf = used && !inner && k.count >= 1;   // else an empty slot (unused or failed a check)
        g.child[j] = leaf ? k.start : 0;                     // inner: assign writes it
        g.count[j] = inner ? 0 : (leaf ? k.count : -1);
        kids_out[s] = inner ? k.entry : -1;
        m += inner ? 1u : 0u;
    }
    return m;
}

// Child indices of wide node w: its inner children are next_base + off + 0,
// 1, ... in slot order, and entries off, off + 1, ... of the next level's
// list.  False when an index reaches cap (the links do not form a tree).
template <int W>
__device__ __forceinline__ bool assign_one(int64_t w, const int32_t* __restrict__ kids, int64_t off,
                                           int64_t next_base, int64_t cap, crt_wide_group* __restrict__ groups,
                                           int32_t* __restrict__ next_entries, int32_t* __restrict__ next_wides,
                                           unsigned* __restrict__ err) {
    int64_t r = 0;
    for (int s = 0; s < W; ++s) {
        const int32_t e = kids[s];
        if (e < 0) continue;
        const int64_t idx = next_base + off + r;
        if (idx >= cap) {   // nothing past the outputs is written
            atomicOr(err, kErrCap);
            return false;
        }
        groups[w * (W / 4) + s / 4].child[s % 4] = int32_t(idx);
        next_entries[off + r] = e;
        next_wides[off + r] = int32_t(idx);
        ++r;
    }
    return true;
}

// A root entry whose two children are leaves with two adjacent triangles in
// all is one 2-triangle leaf under a root of one slot (crt_wide_build's
// rule).  True then, with that leaf's box and range in `k`.
__device__ __forceinline__ bool root_one_slot(const crt_stack_node& root, Kid& k) {
    const crt_bounds l = root.left, r = root.right;
    const int32_t a = __float_as_int(l.min[3]), c = __float_as_int(r.min[3]);
    const int32_t sa = a >> 4, na = a & 0xF, sc = c >> 4, nc = c & 0xF;
    for (int q = 0; q < 3; ++q) {
        k.lo[q] = fminf(l.min[q], r.min[q]);
        k.hi[q] = fmaxf(l.max[q], r.max[q]);
    }
    k.entry = -1;
    k.start = min(sa, sc);
    k.count = na + nc;
    return a != -1 && c != -1 && na + nc <= 2 && max(sa + na, sc + nc) - min(sa, sc) == na + nc;
}

// The one-slot root: slot 0 the leaf (written empty when its range failed
// leaf_ok), every other slot empty.
template <int W>
__device__ __forceinline__ void write_one_slot_root(crt_wide_group* __restrict__ groups, const Kid& k, bool ok) {
    for (int q = 0; q < W / 4; ++q) {
        crt_wide_group& g = groups[q];
        for (int j = 0; j < 4; ++j) {
            const bool s0 = q == 0 && j == 0;
            g.lo_x[j] = s0 ? k.lo[0] : 0.0f;
            g.lo_y[j] = s0 ? k.lo[1] : 0.0f;
            g.lo_z[j] = s0 ? k.lo[2] : 0.0f;
            g.hi_x[j] = s0 ? k.hi[0] : 0.0f;
            g.hi_y[j] = s0 ? k.hi[1] : 0.0f;
            g.hi_z[j] = s0 ? k.hi[2] : 0.0f;
            g.child[j] = s0 && ok ? k.start : 0;
            g.count[j] = s0 && ok ? k.count : -1;
        }
    }
}

}  // namespace wg
}  // namespace crt
