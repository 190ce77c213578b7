// Device collapse of one object's stack-layout BVH2 into the wide layout
// (crt_wide_build_gpu), for trees built on the device every frame
// (crt_build_bvh_gpu) that the wide walks should trace without a host round
// trip.
//
// It is crt_wide_build's collapse (wide_build.cpp) with 2-triangle leaves,
// level by level.  Each wide node of a level expands its BVH2 node as the host
// does: starting from the node's two children, the inner child of largest
// half surface area (the first one on a tie) is replaced by its own two
// children until the node holds `width` of them; a leaf, or an inner node of
// two adjacent 1-triangle leaves, stays one slot.  So every wide node has the
// host's slots, boxes, leaf ranges and slot order.  Nodes are numbered
// breadth-first, where the host numbers them depth-first as it allocates;
// walks visit the same boxes in the same order, so hit records, visit and
// test counts are identical.
//
// Nothing is read back between levels: the host launches a fixed number of
// levels (two launches each: expand, then assign), the level sizes live on the
// device, and a level past the tree's depth returns at once
// (crt_wide_build_gpu_async).  crt_wide_build_gpu runs the same launches and
// reads the node count and status back once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "candela_rt.h"
#include "crt_internal.h"
#include "wide_collapse.h"

namespace crt {
namespace wg {

constexpr int kBlock = 256;

// Exclusive prefix of v over the workgroup (kBlock threads) and the total.
__device__ __forceinline__ uint32_t block_excl(uint32_t v, uint32_t* wave_sums, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == 63) wave_sums[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        before += w < wave ? wave_sums[w] : 0u;
        total += wave_sums[w];
    }
    return before + incl - v;
}

// ---- the asynchronous collapse (crt_wide_build_gpu_async) ----------------
//
// The host launches a fixed number of levels and reads nothing back, so the
// whole collapse can be captured into a graph.  Level state lives on the
// device, two int64 per level l: lvl[2l] is the number of wide nodes numbered
// through level l (the first index of level l + 1) and lvl[2l + 1] the size of
// level l.  Each level launches kGrid-or-fewer workgroups over contiguous
// chunks of the level (so the breadth-first numbering stays the host loop's),
// and a level of size 0 returns at once.

// Bit of *err: the tree has more levels than were launched.
constexpr unsigned kErrDepth = 8u;
static_assert(kErrLink == CRT_WIDE_GPU_BAD_LINK && kErrCap == CRT_WIDE_GPU_NOT_TREE &&
                  kErrVertex == CRT_WIDE_GPU_BAD_VERTEX && kErrDepth == CRT_WIDE_GPU_TOO_DEEP &&
                  kErrLeaf == CRT_WIDE_GPU_BAD_LEAF,
              "status bits are the header's");
constexpr int kGridMax = 2048;

// Exclusive prefix of v over the workgroup and the total, then a barrier, so
// that the call can repeat in a loop.
__device__ __forceinline__ uint32_t block_excl_loop(uint32_t v, uint32_t* wave_sums, uint32_t& total) {
    const uint32_t r = block_excl(v, wave_sums, total);
    __syncthreads();
    return r;
}

// Sum over the workgroup (kBlock threads); every thread gets it.
__device__ __forceinline__ uint64_t block_sum64(uint64_t v, uint64_t* wave_sums) {
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63) == 0) wave_sums[threadIdx.x >> 6] = v;
    __syncthreads();
    uint64_t t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += wave_sums[w];
    __syncthreads();
    return t;
}

// Items [begin, end) of a level of nf wide nodes for this workgroup of a grid
// of `grid`; chunks are whole multiples of kBlock.
__device__ __forceinline__ void chunk_of(int64_t nf, int32_t grid, int64_t& begin, int64_t& end) {
    const int64_t pieces = (nf + kBlock - 1) / kBlock;
    const int64_t per = (pieces + grid - 1) / grid * kBlock;
    begin = int64_t(blockIdx.x) * per;
    end = min(nf, begin + per);
}

// Level 0 and the error word.  A root entry that holds two adjacent
// 1-triangle leaves is one 2-triangle leaf under a root of one slot
// (crt_wide_build's rule): this kernel writes that root, and level 0 is then
// empty.  Otherwise level 0 is the root entry.
template <int W>
__global__ void __launch_bounds__(64) root_async(const crt_stack_node* __restrict__ nodes, int64_t ntris,
                                                 int32_t* __restrict__ entry, int32_t* __restrict__ wide,
                                                 int64_t* __restrict__ lvl, crt_wide_group* __restrict__ groups,
                                                 unsigned* __restrict__ err) {
    if (threadIdx.x != 0) return;
    *err = 0u;
    entry[0] = 0;
    wide[0] = 0;
    Kid k;
    const bool one = root_one_slot(nodes[0], k);
    if (one) {
        const bool ok = leaf_ok(k.start, k.count, ntris);
        if (!ok) *err = kErrLeaf;   // finish_async leaves the root empty
        write_one_slot_root<W>(groups, k, ok);
    }
    lvl[0] = 1;
    lvl[1] = one ? 0 : 1;
}

// expand over the chunks of level `level`; totals[block] for every block of
// the grid when the level is not empty.
template <int W>
__global__ void __launch_bounds__(kBlock) expand_async(const crt_stack_node* __restrict__ nodes, int64_t node_count,
                                                       int64_t ntris, const int64_t* __restrict__ lvl, int32_t level, int32_t grid,
                                                       const int32_t* __restrict__ entries,
                                                       const int32_t* __restrict__ wides,
                                                       crt_wide_group* __restrict__ groups,
                                                       int32_t* __restrict__ kids_out, uint32_t* __restrict__ counts,
                                                       uint32_t* __restrict__ totals, unsigned* __restrict__ err) {
    __shared__ uint32_t wave_sums[kBlock / 64];
    const int64_t nf = lvl[2 * level + 1];
    if (nf == 0) return;
    int64_t begin, end;
    chunk_of(nf, grid, begin, end);
    uint32_t run = 0;
    for (int64_t base = begin; base < end; base += kBlock) {
        const int64_t item = base + threadIdx.x;
        uint32_t m = 0;
        if (item < end) {
            m = expand_one<W>(nodes, node_count, ntris, entries[item], wides[item], groups, kids_out + item * W,
                              err);
            counts[item] = m;
        }
        uint32_t total;
        (void)block_excl_loop(m, wave_sums, total);
        run += total;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = run;
}

// assign over the same chunks; block 0 also writes level + 1's state.  A tree
// whose next level would pass cap stops there (kErrCap).
template <int W>
__global__ void __launch_bounds__(kBlock) assign_async(int64_t* __restrict__ lvl, int32_t level, int32_t grid,
                                                       int64_t cap, const int32_t* __restrict__ wides,
                                                       const int32_t* __restrict__ kids,
                                                       const uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ totals,
                                                       crt_wide_group* __restrict__ groups,
                                                       int32_t* __restrict__ next_entries,
                                                       int32_t* __restrict__ next_wides, unsigned* __restrict__ err) {
    __shared__ uint32_t wave_sums[kBlock / 64];
    __shared__ uint64_t sums[kBlock / 64];
    const int64_t nf = lvl[2 * level + 1], next_base = lvl[2 * level];
    if (nf == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            lvl[2 * level + 2] = next_base;
            lvl[2 * level + 3] = 0;
        }
        return;
    }
    uint64_t before = 0, all = 0;
    for (int32_t j = int32_t(threadIdx.x); j < grid; j += kBlock) {
        const uint64_t v = totals[j];
        all += v;
        before += j < int32_t(blockIdx.x) ? v : 0u;
    }
    before = block_sum64(before, sums);
    all = block_sum64(all, sums);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const bool over = int64_t(next_base + all) > cap;
        if (over) atomicOr(err, kErrCap);
        lvl[2 * level + 2] = over ? next_base : next_base + int64_t(all);
        lvl[2 * level + 3] = over ? 0 : int64_t(all);
    }
    int64_t begin, end;
    chunk_of(nf, grid, begin, end);
    int64_t run = int64_t(before);
    for (int64_t base = begin; base < end; base += kBlock) {
        const int64_t item = base + threadIdx.x;
        const uint32_t m = item < end ? counts[item] : 0u;
        uint32_t total;
        const uint32_t local = block_excl_loop(m, wave_sums, total);
        if (item < end)
            (void)assign_one<W>(wides[item], kids + item * W, run + local, next_base, cap, groups, next_entries,
                                next_wides, err);
        run += total;
    }
}

// The node count, the number of levels the tree needs (the first empty
// level; max_levels when none was) and the status word for the caller.  A
// layout with any status bit is incomplete (levels never expanded hold stale
// memory, a failed slot reads as empty): its root is rewritten with every
// slot empty, so a trace that runs after it, in the same graph say, ends at
// the root with no hit.
template <int W>
__global__ void __launch_bounds__(64) finish_async(const int64_t* __restrict__ lvl, int32_t levels,
                                                   const unsigned* __restrict__ err, int64_t* __restrict__ wide_nodes,
                                                   uint32_t* __restrict__ status, crt_wide_group* __restrict__ groups) {
    const uint32_t st = *err | (lvl[2 * levels + 1] > 0 ? kErrDepth : 0u);
    if (st != 0u) {
        // 64 lanes: W / 4 groups of 8 fields of 4 words (32 words each)
        for (int k = int(threadIdx.x); k < (W / 4) * 32; k += 64) {
            int32_t* g = reinterpret_cast<int32_t*>(groups + k / 32);
            const int word = k % 32;
            g[word] = word >= 28 ? -1 : 0;   // count[4] is the last field
        }
    }
    if (threadIdx.x != 0) return;
    int32_t used = levels;
    for (int32_t l = 0; l < levels; ++l)
        if (lvl[2 * l + 1] == 0) {
            used = l;
            break;
        }
    wide_nodes[0] = lvl[2 * levels];
    wide_nodes[1] = used;
    *status = st;
}

// Triangle records in leaf order (crt_wide_build's).
__global__ void __launch_bounds__(kBlock) records(const crt_triangle* __restrict__ tris,
                                                  const crt_vertex* __restrict__ verts, int64_t nverts, int64_t n,
                                                  crt_tri_record* __restrict__ out, unsigned* __restrict__ err) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    crt_triangle t = tris[i];
    for (int k = 0; k < 3; ++k) {
        if (t.packed[k] < 0 || int64_t(t.packed[k]) >= nverts) {
            atomicOr(err, kErrVertex);
            t.packed[k] = 0;
        }
    }
    crt_tri_record r;
    const float4 a = *reinterpret_cast<const float4*>(verts[t.packed[0]].position);
    const float4 b = *reinterpret_cast<const float4*>(verts[t.packed[1]].position);
    const float4 c = *reinterpret_cast<const float4*>(verts[t.packed[2]].position);
    r.v0[0] = a.x, r.v0[1] = a.y, r.v0[2] = a.z;
    r.v1[0] = b.x, r.v1[1] = b.y, r.v1[2] = b.z;
    r.v2[0] = c.x, r.v2[1] = c.y, r.v2[2] = c.z;
    r.tri = int32_t(i);
    r.mesh = t.packed[3];
    r.pad = 0;
    out[i] = r;
}

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct Layout {
    size_t entries[2], wides[2], kids, counts, totals, err, lvl, out, bytes;
};

// Levels hold at most `cap` wide nodes (a valid tree has fewer wide nodes
// than triangles); the level state has room for `levels` levels.
inline Layout layout(int64_t cap, int32_t width, int32_t levels) {
    const size_t n = size_t(cap), nb = size_t((cap + kBlock - 1) / kBlock);
    Layout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = align256(off + bytes);
        return at;
    };
    L.entries[0] = take(n * 4);
    L.entries[1] = take(n * 4);
    L.wides[0] = take(n * 4);
    L.wides[1] = take(n * 4);
    L.kids = take(n * size_t(width) * 4);
    L.counts = take(n * 4);
    L.totals = take(nb * 4);
    L.err = take(4);
    L.lvl = take(size_t(2) * size_t(levels + 1) * 8);
    L.out = take(3 * 8);   // crt_wide_build_gpu's node count, level count and status
    L.bytes = off;
    return L;
}

template <int W>
void collapse_async(const crt_stack_node* nodes, int64_t node_count, const crt_triangle* tris,
                    const crt_vertex* verts, int64_t nverts, int64_t cap, int32_t levels, crt_wide_group* groups,
                    crt_tri_record* recs, int64_t* wide_nodes, uint32_t* status, char* ws, const Layout& L,
                    hipStream_t st) {
    int32_t* entries[2] = {reinterpret_cast<int32_t*>(ws + L.entries[0]), reinterpret_cast<int32_t*>(ws + L.entries[1])};
    int32_t* wides[2] = {reinterpret_cast<int32_t*>(ws + L.wides[0]), reinterpret_cast<int32_t*>(ws + L.wides[1])};
    auto* kids = reinterpret_cast<int32_t*>(ws + L.kids);
    auto* counts = reinterpret_cast<uint32_t*>(ws + L.counts);
    auto* totals = reinterpret_cast<uint32_t*>(ws + L.totals);
    auto* lvl = reinterpret_cast<int64_t*>(ws + L.lvl);
    auto* err = reinterpret_cast<unsigned*>(ws + L.err);
    const int32_t grid = int32_t(std::min<int64_t>(kGridMax, (cap + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(root_async<W>, dim3(1), dim3(64), 0, st, nodes, cap, entries[0], wides[0], lvl, groups, err);
    hipLaunchKernelGGL(records, dim3(unsigned((cap + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, tris, verts, nverts,
                       cap, recs, err);
    for (int32_t l = 0; l < levels; ++l) {
        const int cur = l & 1;
        hipLaunchKernelGGL(expand_async<W>, dim3(grid), dim3(kBlock), 0, st, nodes, node_count, cap, lvl, l, grid,
                           entries[cur], wides[cur], groups, kids, counts, totals, err);
        hipLaunchKernelGGL(assign_async<W>, dim3(grid), dim3(kBlock), 0, st, lvl, l, grid, cap, wides[cur], kids, counts,
                           totals, groups, entries[1 - cur], wides[1 - cur], err);
    }
    hipLaunchKernelGGL(finish_async<W>, dim3(1), dim3(64), 0, st, lvl, levels, err, wide_nodes, status, groups);
}

}  // namespace wg
}  // namespace crt

extern "C" {

size_t crt_wide_build_gpu_workspace(int64_t ntris, int32_t width) {
    if (ntris < 2 || (width != 4 && width != 8)) return 0;
    return crt::wg::layout(ntris, width, CRT_WIDE_GPU_MAX_LEVELS).bytes;
}

// The asynchronous collapse with kFirstLevels levels, and again with every
// level CRT_WIDE_GPU_MAX_LEVELS allows when the tree is deeper; one readback
// of the node count and status per run.
int crt_wide_build_gpu(const crt_stack_node* nodes, int64_t node_count, const crt_triangle* triangles, int64_t ntris,
                       const crt_vertex* vertices, int64_t nverts, int32_t width, crt_wide_group* groups,
                       crt_tri_record* records, int64_t* wide_nodes, void* workspace, size_t workspace_bytes,
                       void* stream) {
    using namespace crt::wg;
    constexpr int32_t kFirstLevels = 32;
    if (!nodes || !triangles || !vertices || !groups || !records || !wide_nodes || !workspace)
        return crt::set_error(CRT_E_INVALID, "crt_wide_build_gpu: null argument");
    if (width != 4 && width != 8) return crt::set_error(CRT_E_INVALID, "crt_wide_build_gpu: width must be 4 or 8");
    if (ntris < 2 || node_count < 1 || nverts < 1 || ntris > 0x7FFFFFFFll)
        return crt::set_error(CRT_E_INVALID, "crt_wide_build_gpu: needs 2 to 2^31 - 1 triangles and a root entry");
    const Layout L = layout(ntris, width, CRT_WIDE_GPU_MAX_LEVELS);
    if (workspace_bytes < L.bytes || reinterpret_cast<uintptr_t>(workspace) % 256 != 0)
        return crt::set_error(CRT_E_INVALID, "crt_wide_build_gpu: workspace must be 256-byte aligned and hold "
                                             "crt_wide_build_gpu_workspace(ntris, width) = " +
                                                 std::to_string(L.bytes) + " bytes");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    char* ws = static_cast<char*>(workspace);
    auto* out = reinterpret_cast<int64_t*>(ws + L.out);
    auto* status = reinterpret_cast<uint32_t*>(out + 2);
    int64_t host_out[3] = {0, 0, 0};
    for (const int32_t levels : {kFirstLevels, int32_t(CRT_WIDE_GPU_MAX_LEVELS)}) {
        if (width == 4)
            collapse_async<4>(nodes, node_count, triangles, vertices, nverts, ntris, levels, groups, records, out,
                              status, ws, L, st);
        else
            collapse_async<8>(nodes, node_count, triangles, vertices, nverts, ntris, levels, groups, records, out,
                              status, ws, L, st);
        if (hipMemcpyAsync(host_out, out, sizeof(host_out), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return crt::set_error(CRT_E_HIP, std::string("crt_wide_build_gpu: ") + hipGetErrorString(hipGetLastError()));
        if (!(uint32_t(host_out[2]) & kErrDepth)) break;
    }
    const uint32_t bad = uint32_t(host_out[2]);
    if (bad & kErrVertex)
        return crt::set_error(CRT_E_INVALID, "crt_wide_build_gpu: a vertex index leaves the vertex array");
    if (bad & kErrLink) return crt::set_error(CRT_E_INVALID, "crt_wide_build_gpu: a stack link leaves the node array");
    if (bad & kErrLeaf)
        return crt::set_error(CRT_E_INVALID, "crt_wide_build_gpu: a leaf range is empty or leaves [0, ntris) (the "
                                             "tree must be built with triangle offset 0)");
    if (bad & kErrCap)
        return crt::set_error(CRT_E_INVALID,
                              "crt_wide_build_gpu: more wide nodes than triangles (the links do not form a tree)");
    if (bad & kErrDepth)
        return crt::set_error(CRT_E_INVALID, "crt_wide_build_gpu: the tree has more than " +
                                                 std::to_string(CRT_WIDE_GPU_MAX_LEVELS) + " wide levels");
    *wide_nodes = host_out[0];
    return CRT_OK;
}

size_t crt_wide_build_gpu_async_workspace(int64_t ntris, int32_t width, int32_t max_levels) {
    if (ntris < 2 || (width != 4 && width != 8) || max_levels < 1 || max_levels > CRT_WIDE_GPU_MAX_LEVELS) return 0;
    return crt::wg::layout(ntris, width, max_levels).bytes;
}

int crt_wide_build_gpu_async(const crt_stack_node* nodes, int64_t node_count, const crt_triangle* triangles,
                             int64_t ntris, const crt_vertex* vertices, int64_t nverts, int32_t width,
                             int32_t max_levels, crt_wide_group* groups, crt_tri_record* records,
                             int64_t* wide_nodes, uint32_t* status, void* workspace, size_t workspace_bytes,
                             void* stream) {
    using namespace crt::wg;
    if (!nodes || !triangles || !vertices || !groups || !records || !wide_nodes || !status || !workspace)
        return crt::set_error(CRT_E_INVALID, "crt_wide_build_gpu_async: null argument");
    if (width != 4 && width != 8)
        return crt::set_error(CRT_E_INVALID, "crt_wide_build_gpu_async: width must be 4 or 8");
    if (max_levels < 1 || max_levels > CRT_WIDE_GPU_MAX_LEVELS)
        return crt::set_error(CRT_E_INVALID, "crt_wide_build_gpu_async: max_levels must be 1 to " +
                                                 std::to_string(CRT_WIDE_GPU_MAX_LEVELS));
    if (ntris < 2 || node_count < 1 || nverts < 1 || ntris > 0x7FFFFFFFll)
        return crt::set_error(CRT_E_INVALID,
                              "crt_wide_build_gpu_async: needs 2 to 2^31 - 1 triangles and a root entry");
    const Layout L = layout(ntris, width, max_levels);
    if (workspace_bytes < L.bytes || reinterpret_cast<uintptr_t>(workspace) % 256 != 0)
        return crt::set_error(CRT_E_INVALID, "crt_wide_build_gpu_async: workspace must be 256-byte aligned and hold "
                                             "crt_wide_build_gpu_async_workspace(ntris, width, max_levels) = " +
                                                 std::to_string(L.bytes) + " bytes");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    char* ws = static_cast<char*>(workspace);
    if (width == 4)
        collapse_async<4>(nodes, node_count, triangles, vertices, nverts, ntris, max_levels, groups, records,
                          wide_nodes, status, ws, L, st);
    else
        collapse_async<8>(nodes, node_count, triangles, vertices, nverts, ntris, max_levels, groups, records,
                          wide_nodes, status, ws, L, st);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return crt::set_error(CRT_E_HIP, std::string("crt_wide_build_gpu_async: ") + hipGetErrorString(e));
    return CRT_OK;
}

}  // extern "C"
