// Batched traversal: many scenes and ray batches in one persistent launch.
//
// The plain path traces one (scene, query) batch per launch; a frame with many
// small batches then pays a launch and a draining tail per batch, and a batch
// of a few thousand rays cannot fill 256 CUs at all.  Here every batch is cut
// into 64-ray tiles, one tile per wave, and device-filling waves claim tiles
// from a tile table that spans all batches (optionally heaviest tiles first,
// so the launch ends on short tiles) through the sharded queue of queue.h:
// shard k holds tiles k, k + 16, ..., so every shard keeps the table's order
// and no counter sees more than a sixteenth of the claims.  A wave traces its
// tile with the same wave routine as the plain stackless kernels
// (wave_walk.h), so every hit record is bit-identical to crt_trace on the same
// batch.  The launch is a queue-reset kernel plus the batched kernel -- no
// memset and no allocation -- so it can be captured once into a hipGraph and
// replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "candela_rt.h"
#include "crt_internal.h"
#include "queue.h"
#include "trace_launch.h"
#include "wave_walk.h"
#include "wide_walk.h"
#include "spec_bvh2.h"
#include "spec_walk.h"
#include "thread_walk.h"

namespace crt {

struct BatchDesc {
    SceneArgs scene;
    const crt_ray* rays;
    crt_hit* hits;
    uint32_t* tri_tests;
    int64_t count;
    int32_t query;
    uint32_t flags;
};

namespace pdev {

// Pointers read from a descriptor in memory are generic to the compiler, so
// every access through them would be a flat instruction, which also counts
// against the LDS wait counter.  All of them point to device memory: rebuilt
// from their address as global-segment pointers, they give global loads and
// stores (a plain cast to the global segment and back would be folded away).
template <class T>
__device__ __forceinline__ T* global_ptr(T* p) {
    return (T*)(__attribute__((address_space(1))) T*)reinterpret_cast<uintptr_t>(p);
}

__device__ __forceinline__ BatchDesc load_desc(const BatchDesc* __restrict__ descs, int32_t b) {
    BatchDesc d = descs[b];
    d.scene.nodes = global_ptr(d.scene.nodes);
    d.scene.tris = global_ptr(d.scene.tris);
    d.scene.verts = global_ptr(d.scene.verts);
    d.scene.entities = global_ptr(d.scene.entities);
    d.scene.tri_records = global_ptr(d.scene.tri_records);
    d.rays = global_ptr(d.rays);
    d.hits = global_ptr(d.hits);
    d.tri_tests = global_ptr(d.tri_tests);
    return d;
}

template <int INNER, int LEAF_BATCH, int K>
__device__ __forceinline__ void stackless_body(const BatchDesc* __restrict__ descs, const int2* __restrict__ tiles,
                                               int32_t ntiles, int32_t per_claim,
                                               unsigned long long* __restrict__ counters, int32_t max_iter,
                                               int32_t shards) {
    // one walk for both queries (the query is a per-tile value): two inlined
    // walks would need the sum of their registers
    WaveWalk<false, INNER, LEAF_BATCH, K, true> w;
    ShardQueue q;
    q.init(counters, ntiles, true, shards);
    while (true) {
        int32_t got = 0;
        const int64_t base = q.claim(per_claim, got);
        if (got == 0) break;
        for (int32_t k = 0; k < got; ++k) {
            const int32_t t = __builtin_amdgcn_readfirstlane(int32_t(q.at(base + k)));
            const int2 tile = tiles[t];
            const int32_t b = __builtin_amdgcn_readfirstlane(tile.x);
            const int32_t first = __builtin_amdgcn_readfirstlane(tile.y);
            const BatchDesc d = load_desc(descs, b);
            const int64_t i = int64_t(first) + __lane_id();
            const int64_t slot = i < d.count ? i : -1;
            w.dyn_any = d.query == CRT_QUERY_ANY;
            w.run(d.scene, d.rays, d.hits, slot, d.flags, max_iter, d.tri_tests);
        }
    }
}

#define CRT_BATCH_ARGS                                                                                          \
    const BatchDesc *__restrict__ descs, const int2 *__restrict__ tiles, int32_t ntiles, int32_t per_claim,      \
        unsigned long long *__restrict__ counters, int32_t max_iter, int32_t shards

template <int INNER, int LEAF_BATCH, int K, int WPG>
__global__ void __launch_bounds__(64 * WPG) batched_stackless(CRT_BATCH_ARGS) {
    stackless_body<INNER, LEAF_BATCH, K>(descs, tiles, ntiles, per_claim, counters, max_iter, shards);
}

// The same kernels compiled for at least OCC waves per SIMD (register budget
// 512 / OCC), trading spills for occupancy.
template <int INNER, int LEAF_BATCH, int K, int WPG, int OCC>
__global__ void __launch_bounds__(64 * WPG) __attribute__((amdgpu_waves_per_eu(OCC, 8)))
batched_stackless_occ(CRT_BATCH_ARGS) {
    stackless_body<INNER, LEAF_BATCH, K>(descs, tiles, ntiles, per_claim, counters, max_iter, shards);
}

// The same tile queue over the wide layouts (BVH4 / BVH8): each claimed tile
// runs the wave-scheduled wide walk (wide_walk.h), so every hit record equals
// crt_trace's wide wave kernel on the same batch.
template <int W, int LDS_DEPTH, int INNER, int LEAF_BATCH, int WPG>
__device__ __forceinline__ void wide_body(const BatchDesc* __restrict__ descs, const int2* __restrict__ tiles,
                                          int32_t ntiles, int32_t per_claim, unsigned long long* __restrict__ counters,
                                          int32_t max_iter, int32_t shards, int32_t* lds) {
    int32_t spill_node[wide::kSpill];
    float spill_t[wide::kSpill];
    wide::Stack<LDS_DEPTH> st;
    st.lds = lds + threadIdx.x;
    st.stride = 64 * WPG;
    st.spill_node = spill_node;
    st.spill_t = spill_t;
    ShardQueue q;
    q.init(counters, ntiles, true, shards);
    while (true) {
        int32_t got = 0;
        const int64_t base = q.claim(per_claim, got);
        if (got == 0) break;
        for (int32_t k = 0; k < got; ++k) {
            const int32_t t = __builtin_amdgcn_readfirstlane(int32_t(q.at(base + k)));
            const int2 tile = tiles[t];
            const int32_t b = __builtin_amdgcn_readfirstlane(tile.x);
            const int32_t first = __builtin_amdgcn_readfirstlane(tile.y);
            const BatchDesc d = load_desc(descs, b);
            const int64_t i = int64_t(first) + __lane_id();
            const int64_t slot = i < d.count ? i : -1;
            // one walk for both queries (the query is a per-tile value): two
            // inlined walks would need the sum of their registers
            wide::WideWave<false, W, LDS_DEPTH, INNER, LEAF_BATCH, true> w;
            w.c.dyn_any = d.query == CRT_QUERY_ANY;
            w.run(d.scene, d.rays, d.hits, slot, d.flags, max_iter, d.tri_tests, st);
        }
    }
}

template <int W, int LDS_DEPTH, int INNER, int LEAF_BATCH, int WPG>
__global__ void __launch_bounds__(64 * WPG) batched_wide(CRT_BATCH_ARGS) {
    __shared__ int32_t lds[wide::lds_ints<W, LDS_DEPTH>() * 64 * WPG];
    wide_body<W, LDS_DEPTH, INNER, LEAF_BATCH, WPG>(descs, tiles, ntiles, per_claim, counters, max_iter, shards, lds);
}

template <int W, int LDS_DEPTH, int INNER, int LEAF_BATCH, int WPG, int OCC>
__global__ void __launch_bounds__(64 * WPG) __attribute__((amdgpu_waves_per_eu(OCC, 8)))
batched_wide_occ(CRT_BATCH_ARGS) {
    __shared__ int32_t lds[wide::lds_ints<W, LDS_DEPTH>() * 64 * WPG];
    wide_body<W, LDS_DEPTH, INNER, LEAF_BATCH, WPG>(descs, tiles, ntiles, per_claim, counters, max_iter, shards, lds);
}
// The speculative wide walk (spec_walk.h) on the same tile queue: a tile is
// rays [64 k, 64 k + 64) of its batch, exactly one wave of crt_trace's
// speculative grid launch on that batch, so with the same preset every hit
// record, counters included, equals crt_trace's kernel 4.
template <int W, int LDS_DEPTH, int INNER, int Q, int WPG>
struct SpecGroupLds {
    int32_t cols[spec::lds_ints<LDS_DEPTH, Q>() * 64 * WPG];
    spec::WaveLds<Q> waves[WPG];
};

template <int W, int LDS_DEPTH, int INNER, int Q, int WPG>
__device__ __forceinline__ void spec_body(const BatchDesc* __restrict__ descs, const int2* __restrict__ tiles,
                                          int32_t ntiles, int32_t per_claim, unsigned long long* __restrict__ counters,
                                          int32_t max_iter, int32_t shards,
                                          SpecGroupLds<W, LDS_DEPTH, INNER, Q, WPG>& lds) {
    int32_t spill_node[wide::kSpill];
    float spill_t[wide::kSpill];
    wide::Stack<LDS_DEPTH> st;
    st.lds = lds.cols + threadIdx.x;
    st.stride = 64 * WPG;
    st.spill_node = spill_node;
    st.spill_t = spill_t;
    spec::WaveLds<Q>& wl = lds.waves[threadIdx.x >> 6];
    ShardQueue q;
    q.init(counters, ntiles, true, shards);
    while (true) {
        int32_t got = 0;
        const int64_t base = q.claim(per_claim, got);
        if (got == 0) break;
        for (int32_t k = 0; k < got; ++k) {
            const int32_t t = __builtin_amdgcn_readfirstlane(int32_t(q.at(base + k)));
            const int2 tile = tiles[t];
            const int32_t b = __builtin_amdgcn_readfirstlane(tile.x);
            const int32_t first = __builtin_amdgcn_readfirstlane(tile.y);
            const BatchDesc d = load_desc(descs, b);
            const int64_t i = int64_t(first) + __lane_id();
            const int64_t slot = i < d.count ? i : -1;
            // one walk for both queries (the query is a per-tile value)
            spec::SpecWave<false, W, LDS_DEPTH, INNER, Q, true> w;
            w.dyn_any = d.query == CRT_QUERY_ANY;
            w.run(d.scene, d.rays, d.hits, slot, d.flags, max_iter, d.tri_tests, st, wl);
        }
    }
}

template <int W, int LDS_DEPTH, int INNER, int Q, int WPG, int OCC>
__global__ void __launch_bounds__(64 * WPG) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1, 8)))
batched_spec(CRT_BATCH_ARGS) {
    __shared__ SpecGroupLds<W, LDS_DEPTH, INNER, Q, WPG> lds;
    spec_body<W, LDS_DEPTH, INNER, Q, WPG>(descs, tiles, ntiles, per_claim, counters, max_iter, shards, lds);
}

// The speculative walks of the two BVH2 layouts (spec_bvh2.h; D = 0 the
// stackless layout, D > 0 the stack layout with D LDS stack entries) on the
// same tile queue: a tile is one wave of crt_trace's speculative grid launch
// on its batch, so with the same preset every hit record, counters included,
// equals crt_trace's kernel 4.
template <int D, int Q, int WPG>
struct Bvh2SpecLds {
    int32_t queue[Q * 64 * WPG];
    int32_t stack[(D > 0 ? D : 1) * 64 * WPG];
    spec::WaveLds<Q> waves[WPG];
};

template <int D, int INNER, int Q, int WPG>
__device__ __forceinline__ void bvh2_spec_body(const BatchDesc* __restrict__ descs, const int2* __restrict__ tiles,
                                               int32_t ntiles, int32_t per_claim,
                                               unsigned long long* __restrict__ counters, int32_t max_iter,
                                               int32_t shards, Bvh2SpecLds<D, Q, WPG>& lds) {
    constexpr int kCols = 64 * WPG;
    int32_t spill[D > 0 ? sdev::spill_entries<D>() : 1];
    spec::WaveLds<Q>& wl = lds.waves[threadIdx.x >> 6];
    ShardQueue q;
    q.init(counters, ntiles, true, shards);
    while (true) {
        int32_t got = 0;
        const int64_t base = q.claim(per_claim, got);
        if (got == 0) break;
        for (int32_t k = 0; k < got; ++k) {
            const int32_t t = __builtin_amdgcn_readfirstlane(int32_t(q.at(base + k)));
            const int2 tile = tiles[t];
            const int32_t b = __builtin_amdgcn_readfirstlane(tile.x);
            const int32_t first = __builtin_amdgcn_readfirstlane(tile.y);
            const BatchDesc d = load_desc(descs, b);
            const int64_t i = int64_t(first) + __lane_id();
            const int64_t slot = i < d.count ? i : -1;
            // one walk for both queries (the query is a per-tile value)
            sbvh2::SpecBvh2Wave<false, D, INNER, Q, true, kCols> w;
            w.dyn_any = d.query == CRT_QUERY_ANY;
            w.q = (LdsInt*)(lds.queue + threadIdx.x);
            if constexpr (D > 0) w.st = sdev::NodeStack<D>{lds.stack + threadIdx.x, kCols, spill};
            w.run(d.scene, d.rays, d.hits, slot, d.flags, max_iter, d.tri_tests, wl);
        }
    }
}

template <int D, int INNER, int Q, int WPG>
__global__ void __launch_bounds__(64 * WPG) batched_bvh2_spec(CRT_BATCH_ARGS) {
    __shared__ Bvh2SpecLds<D, Q, WPG> lds;
    bvh2_spec_body<D, INNER, Q, WPG>(descs, tiles, ntiles, per_claim, counters, max_iter, shards, lds);
}

// The stack format (FlattenStackBVH): each claimed tile runs the one-ray-per-
// lane stack walk of crt_trace's thread kernel (thread_walk.h), its first
// LDS_DEPTH entries in LDS; the walk's results do not depend on LDS_DEPTH, so
// every hit record equals crt_trace's.
template <int LDS_DEPTH, int WPG>
__device__ __forceinline__ void stack_body(const BatchDesc* __restrict__ descs, const int2* __restrict__ tiles,
                                           int32_t ntiles, int32_t per_claim, unsigned long long* __restrict__ counters,
                                           int32_t max_iter, int32_t shards, int32_t* lds) {
    ShardQueue q;
    q.init(counters, ntiles, true, shards);
    while (true) {
        int32_t got = 0;
        const int64_t base = q.claim(per_claim, got);
        if (got == 0) break;
        for (int32_t k = 0; k < got; ++k) {
            const int32_t t = __builtin_amdgcn_readfirstlane(int32_t(q.at(base + k)));
            const int2 tile = tiles[t];
            const int32_t b = __builtin_amdgcn_readfirstlane(tile.x);
            const int32_t first = __builtin_amdgcn_readfirstlane(tile.y);
            const BatchDesc d = load_desc(descs, b);
            const int64_t i = int64_t(first) + __lane_id();
            if (i >= d.count) continue;
            const crt_ray r = dev::load_ray(d.rays, i);
            uint32_t* tt = d.tri_tests ? d.tri_tests + i : nullptr;
            if (d.query == CRT_QUERY_ANY)
                dev::trace_one<CRT_FORMAT_STACK, true, LDS_DEPTH>(d.scene, r, d.flags, max_iter, d.hits + i, tt,
                                                                 lds + threadIdx.x, 64 * WPG);
            else
                dev::trace_one<CRT_FORMAT_STACK, false, LDS_DEPTH>(d.scene, r, d.flags, max_iter, d.hits + i, tt,
                                                                  lds + threadIdx.x, 64 * WPG);
        }
    }
}

template <int LDS_DEPTH, int WPG>
__global__ void __launch_bounds__(64 * WPG) batched_stack(CRT_BATCH_ARGS) {
    __shared__ int32_t lds[LDS_DEPTH * 64 * WPG];
    stack_body<LDS_DEPTH, WPG>(descs, tiles, ntiles, per_claim, counters, max_iter, shards, lds);
}

template <int LDS_DEPTH, int WPG, int OCC>
__global__ void __launch_bounds__(64 * WPG) __attribute__((amdgpu_waves_per_eu(OCC, 8)))
batched_stack_occ(CRT_BATCH_ARGS) {
    __shared__ int32_t lds[LDS_DEPTH * 64 * WPG];
    stack_body<LDS_DEPTH, WPG>(descs, tiles, ntiles, per_claim, counters, max_iter, shards, lds);
}
#undef CRT_BATCH_ARGS

}  // namespace pdev

using BatchedFn = void (*)(const BatchDesc*, const int2*, int32_t, int32_t, unsigned long long*, int32_t, int32_t);

// Speculative BVH2 presets 8..11 = crt_trace's speculative grid presets 0..3
// over the same layout (INNER, queue slots; the stack layout keeps 8 stack
// entries in LDS).
template <int D, int WPG>
static const void* batched_bvh2_spec_symbol(int variant) {
    switch (variant) {
        case 9: return reinterpret_cast<const void*>(&pdev::batched_bvh2_spec<D, 16, 8, WPG>);
        case 10: return reinterpret_cast<const void*>(&pdev::batched_bvh2_spec<D, 16, 12, WPG>);
        case 11: return reinterpret_cast<const void*>(&pdev::batched_bvh2_spec<D, 32, 16, WPG>);
        default: return reinterpret_cast<const void*>(&pdev::batched_bvh2_spec<D, 8, 8, WPG>);
    }
}

template <int WPG>
static const void* batched_symbol(int variant) {
    if (variant >= 8) return batched_bvh2_spec_symbol<0, WPG>(variant);
    switch (variant) {
        case 1: return reinterpret_cast<const void*>(&pdev::batched_stackless<8, 8, 1, WPG>);
        case 2: return reinterpret_cast<const void*>(&pdev::batched_stackless<8, 8, 4, WPG>);
        case 3: return reinterpret_cast<const void*>(&pdev::batched_stackless<16, 16, 2, WPG>);
        case 4: return reinterpret_cast<const void*>(&pdev::batched_stackless<4, 4, 2, WPG>);
        case 5: return reinterpret_cast<const void*>(&pdev::batched_stackless<8, 12, 4, WPG>);
        case 6: return reinterpret_cast<const void*>(&pdev::batched_stackless_occ<8, 8, 2, WPG, 5>);
        case 7: return reinterpret_cast<const void*>(&pdev::batched_stackless_occ<8, 8, 1, WPG, 6>);
        default: return reinterpret_cast<const void*>(&pdev::batched_stackless<8, 8, 2, WPG>);
    }
}

// Wide presets (LDS stack entries, INNER, LEAF_BATCH) -- the same as the
// plain wide wave kernel's (wide.hip).
// Speculative presets 8..11 = crt_trace's speculative grid presets 0, 1, 3
// and 4 (LDS stack entries, INNER, queue slots, OCC); the BVH8 walk takes 4
// more queue slots, as crt_trace's does.
template <int W, int WPG>
static const void* batched_spec_symbol(int variant) {
    constexpr int Q = W == 4 ? 8 : 12;
    switch (variant) {
        case 9: return reinterpret_cast<const void*>(&pdev::batched_spec<W, 8, 8, Q, WPG, 0>);
        case 10: return reinterpret_cast<const void*>(&pdev::batched_spec<W, 8, 4, Q, WPG, 5>);
        case 11: return reinterpret_cast<const void*>(&pdev::batched_spec<W, 8, 8, Q, WPG, 5>);
        default: return reinterpret_cast<const void*>(&pdev::batched_spec<W, 8, 4, Q, WPG, 0>);
    }
}

template <int W, int WPG>
static const void* batched_wide_symbol(int variant) {
    if (variant >= 8) return batched_spec_symbol<W, WPG>(variant);
    switch (variant) {
        case 1: return reinterpret_cast<const void*>(&pdev::batched_wide<W, 8, 8, 16, WPG>);
        case 2: return reinterpret_cast<const void*>(&pdev::batched_wide<W, 8, 2, 8, WPG>);
        case 3: return reinterpret_cast<const void*>(&pdev::batched_wide<W, 16, 4, 16, WPG>);
        case 4: return reinterpret_cast<const void*>(&pdev::batched_wide<W, 8, 4, 32, WPG>);
        case 5: return reinterpret_cast<const void*>(&pdev::batched_wide<W, 8, 16, 48, WPG>);
        case 6: return reinterpret_cast<const void*>(&pdev::batched_wide_occ<W, 8, 8, 16, WPG, 5>);
        case 7: return reinterpret_cast<const void*>(&pdev::batched_wide_occ<W, 8, 4, 16, WPG, 5>);
        default: return reinterpret_cast<const void*>(&pdev::batched_wide<W, 8, 4, 16, WPG>);
    }
}

// Stack presets: LDS stack entries per lane (0..5), then 8 entries compiled
// for at least 5 (6) or 6 (7) waves per SIMD.
template <int WPG>
static const void* batched_stack_symbol(int variant) {
    if (variant >= 8) return batched_bvh2_spec_symbol<8, WPG>(variant);
    switch (variant) {
        case 1: return reinterpret_cast<const void*>(&pdev::batched_stack<16, WPG>);
        case 2: return reinterpret_cast<const void*>(&pdev::batched_stack<32, WPG>);
        case 3: return reinterpret_cast<const void*>(&pdev::batched_stack<4, WPG>);
        case 4: return reinterpret_cast<const void*>(&pdev::batched_stack<12, WPG>);
        case 5: return reinterpret_cast<const void*>(&pdev::batched_stack<24, WPG>);
        case 6: return reinterpret_cast<const void*>(&pdev::batched_stack_occ<8, WPG, 5>);
        case 7: return reinterpret_cast<const void*>(&pdev::batched_stack_occ<8, WPG, 6>);
        default: return reinterpret_cast<const void*>(&pdev::batched_stack<8, WPG>);
    }
}

template <int WPG>
static const void* kernel_for(int format, int variant) {
    if (format == CRT_FORMAT_STACK) return batched_stack_symbol<WPG>(variant);
    if (format == CRT_FORMAT_WIDE4) return batched_wide_symbol<4, WPG>(variant);
    if (format == CRT_FORMAT_WIDE8) return batched_wide_symbol<8, WPG>(variant);
    return batched_symbol<WPG>(variant);
}

static const void* batched_kernel(int format, int variant, int wpg) {
    switch (wpg) {
        case 1: return kernel_for<1>(format, variant);
        case 2: return kernel_for<2>(format, variant);
        default: return kernel_for<4>(format, variant);
    }
}

}  // namespace crt

struct crt_batch_plan {
    std::vector<crt_batch> batches;
    crt_batch_config cfg{};
    crt::BatchDesc* d_descs = nullptr;
    int2* d_tiles = nullptr;
    unsigned long long* d_counters = nullptr;   // one sharded tile queue (queue.h)
    int32_t ntiles = 0;
    int64_t rays = 0;
    int grid = 0;
    const void* kernel = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;

    ~crt_batch_plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (d_descs) (void)hipFree(d_descs);
        if (d_tiles) (void)hipFree(d_tiles);
        if (d_counters) (void)hipFree(d_counters);
    }
};

namespace {

int hip_err(hipError_t e, const char* what) {
    return crt::set_error(CRT_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define PLAN_HIP(call, what)                      \
    do {                                          \
        hipError_t e_ = (call);                   \
        if (e_ != hipSuccess) return hip_err(e_, what); \
    } while (0)

static_assert(sizeof(crt_batch_config) == 28, "crt_batch_config is 7 int32 fields");

int validate_config(const crt_batch_config& c) {
    if (c.variant < 0 || c.variant > 11)
        return crt::set_error(CRT_E_INVALID, "batch config: variant must be 0..11");
    if (c.tiles_per_claim != 1 && c.tiles_per_claim != 2 && c.tiles_per_claim != 4)
        return crt::set_error(CRT_E_INVALID, "batch config: tiles_per_claim must be 1, 2 or 4");
    if (c.waves_per_group != 1 && c.waves_per_group != 2 && c.waves_per_group != 4)
        return crt::set_error(CRT_E_INVALID, "batch config: waves_per_group must be 1, 2 or 4");
    if (c.grid < 0 || c.order < 0 || c.order > 1) return crt::set_error(CRT_E_INVALID, "batch config: bad grid/order");
    if (c.shards != 0 && c.shards != 1 && c.shards != crt::kQueueShards)
        return crt::set_error(CRT_E_INVALID, "batch config: shards must be 0, 1 or 16");
    if (c.child_order != 0 && c.child_order != 1)
        return crt::set_error(CRT_E_INVALID, "batch config: child_order must be 0 or 1");
    return CRT_OK;
}

crt_batch_config default_config() {
    crt_batch_config c{};
    c.variant = 0;
    c.tiles_per_claim = 1;
    c.waves_per_group = 4;
    c.grid = 0;
    c.order = 1;
    return c;
}

// Relative cost of one tile: any-hit walks stop at the first hit; a walk's
// length grows with the depth of the tree.
double tile_cost(const crt_batch& b) {
    const double depth = std::log2(double(std::max<int64_t>(2, b.scene.total_nodes)));
    return (b.query == CRT_QUERY_ANY ? 1.0 : 2.5) * depth;
}

int build_plan(const crt_batch* batches, int nbatches, const crt_batch_config& cfg, crt_batch_plan& p) {
    p.batches.assign(batches, batches + nbatches);
    p.cfg = cfg;
    std::vector<crt::BatchDesc> descs(static_cast<size_t>(nbatches));
    const int format = nbatches > 0 ? batches[0].scene.format : CRT_FORMAT_STACKLESS;
    const bool wide = format == CRT_FORMAT_WIDE4 || format == CRT_FORMAT_WIDE8;
    for (int i = 0; i < nbatches; ++i) {
        const crt_batch& b = batches[i];
        if (b.scene.format != CRT_FORMAT_STACKLESS && b.scene.format != CRT_FORMAT_STACK &&
            b.scene.format != CRT_FORMAT_WIDE4 && b.scene.format != CRT_FORMAT_WIDE8)
            return crt::set_error(CRT_E_INVALID, "batch: unknown node format");
        if (b.scene.format != format)
            return crt::set_error(CRT_E_INVALID, "batch plan: every batch must use the same node format");
        if (b.query != CRT_QUERY_CLOSEST && b.query != CRT_QUERY_ANY)
            return crt::set_error(CRT_E_INVALID, "batch: bad query");
        if (b.count < 0 || (b.count > 0 && (!b.rays || !b.hits)))
            return crt::set_error(CRT_E_INVALID, "batch: bad ray/hit buffers");
        if (b.count > 0 && b.scene.entity_count > 0 && !wide &&
            (!b.scene.nodes || !b.scene.triangles || !b.scene.vertices || !b.scene.entities || b.scene.total_nodes <= 0))
            return crt::set_error(CRT_E_INVALID, "batch: null scene buffer");
        if (b.count > 0 && b.scene.entity_count > 0 && wide &&
            (!b.scene.nodes || !b.scene.tri_records || !b.scene.entities || b.scene.total_nodes <= 0))
            return crt::set_error(CRT_E_INVALID, "batch: null scene buffer (wide layout)");
        if (b.count >= (int64_t(1) << 31)) return crt::set_error(CRT_E_INVALID, "batch: >= 2^31 rays");
        crt::BatchDesc& d = descs[size_t(i)];
        d.scene = crt::scene_args(b.scene);
        d.rays = b.rays;
        d.hits = b.hits;
        d.tri_tests = b.tri_tests;
        d.count = b.count;
        d.query = b.query;
        d.flags = b.flags | (cfg.child_order == 1 && b.query == CRT_QUERY_ANY ? uint32_t(CRT_FLAG_SLOT_ORDER) : 0u);
    }
    std::vector<int> order(static_cast<size_t>(nbatches));
    std::iota(order.begin(), order.end(), 0);
    if (cfg.order == 1)
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return tile_cost(batches[a]) > tile_cost(batches[b]); });
    std::vector<int2> tiles;
    p.rays = 0;
    for (int bi : order) {
        const int64_t n = batches[bi].count;
        p.rays += n;
        for (int64_t first = 0; first < n; first += 64) tiles.push_back(make_int2(bi, int(first)));
    }
    if (tiles.size() >= size_t(1u << 31)) return crt::set_error(CRT_E_INVALID, "batch plan: too many tiles");
    p.ntiles = int32_t(tiles.size());
    PLAN_HIP(hipMalloc(&p.d_descs, std::max<size_t>(1, descs.size()) * sizeof(crt::BatchDesc)), "hipMalloc(descs)");
    PLAN_HIP(hipMalloc(&p.d_tiles, std::max<size_t>(1, tiles.size()) * sizeof(int2)), "hipMalloc(tiles)");
    PLAN_HIP(hipMalloc(&p.d_counters, size_t(crt::kQueueWords) * sizeof(unsigned long long)), "hipMalloc(queue)");
    if (!descs.empty())
        PLAN_HIP(hipMemcpy(p.d_descs, descs.data(), descs.size() * sizeof(crt::BatchDesc), hipMemcpyHostToDevice),
                 "hipMemcpy(descs)");
    if (!tiles.empty())
        PLAN_HIP(hipMemcpy(p.d_tiles, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice),
                 "hipMemcpy(tiles)");
    p.kernel = crt::batched_kernel(format, cfg.variant, cfg.waves_per_group);
    int grid = cfg.grid;
    if (grid == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        PLAN_HIP(hipGetDevice(&dev), "hipGetDevice");
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, p.kernel, 64 * cfg.waves_per_group, 0) !=
                hipSuccess ||
            per_cu <= 0)
            per_cu = 16 / cfg.waves_per_group;
        grid = cus * per_cu;
    }
    // no more waves than tiles
    const int64_t max_groups = (int64_t(p.ntiles) + cfg.waves_per_group - 1) / cfg.waves_per_group;
    p.grid = int(std::max<int64_t>(1, std::min<int64_t>(grid, max_groups)));
    return CRT_OK;
}

int launch_plan(crt_batch_plan& p, hipStream_t stream) {
    if (p.ntiles == 0) return CRT_OK;
    hipLaunchKernelGGL(crt::reset_queue, dim3(1), dim3(64), 0, stream, p.d_counters);
    PLAN_HIP(hipGetLastError(), "batched: reset launch");
    const crt::BatchDesc* descs = p.d_descs;
    const int2* tiles = p.d_tiles;
    int32_t ntiles = p.ntiles;
    int32_t per_claim = p.cfg.tiles_per_claim;
    unsigned long long* counters = p.d_counters;
    int32_t max_iter = crt::kMaxIterations;
    int32_t shards = p.cfg.shards > 1 ? p.cfg.shards : 1;
    void* args[] = {&descs, &tiles, &ntiles, &per_claim, &counters, &max_iter, &shards};
    PLAN_HIP(hipLaunchKernel(p.kernel, dim3(unsigned(p.grid)), dim3(unsigned(64 * p.cfg.waves_per_group)), args, 0,
                             stream),
             "batched: kernel launch");
    return CRT_OK;
}

}  // namespace

extern "C" {

int crt_batch_plan_create(const crt_batch* batches, int nbatches, const crt_batch_config* config,
                          crt_batch_plan** out) {
    if (!out) return crt::set_error(CRT_E_INVALID, "crt_batch_plan_create: out is null");
    *out = nullptr;
    if (nbatches < 0 || (nbatches > 0 && !batches))
        return crt::set_error(CRT_E_INVALID, "crt_batch_plan_create: bad batch list");
    const crt_batch_config cfg = config ? *config : default_config();
    int rc = validate_config(cfg);
    if (rc) return rc;
    crt_batch_plan* p = new (std::nothrow) crt_batch_plan();
    if (!p) return crt::set_error(CRT_E_NOMEM, "crt_batch_plan_create: out of host memory");
    rc = build_plan(batches, nbatches, cfg, *p);
    if (rc) {
        delete p;
        return rc;
    }
    *out = p;
    return CRT_OK;
}

void crt_batch_plan_destroy(crt_batch_plan* plan) { delete plan; }

int crt_batch_plan_launch(crt_batch_plan* plan, void* stream) {
    if (!plan) return crt::set_error(CRT_E_INVALID, "crt_batch_plan_launch: plan is null");
    return launch_plan(*plan, static_cast<hipStream_t>(stream));
}

int crt_batch_plan_launch_graph(crt_batch_plan* plan, void* stream) {
    if (!plan) return crt::set_error(CRT_E_INVALID, "crt_batch_plan_launch_graph: plan is null");
    if (!plan->exec) {
        // Stream capture needs a non-default stream; the graph is then launched
        // on the caller's stream.
        hipStream_t cap = nullptr;
        PLAN_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        int rc = CRT_OK;
        if (e == hipSuccess) {
            rc = launch_plan(*plan, cap);
            hipGraph_t g = nullptr;
            const hipError_t e2 = hipStreamEndCapture(cap, &g);
            if (rc == CRT_OK && e2 != hipSuccess) rc = hip_err(e2, "hipStreamEndCapture");
            if (rc == CRT_OK) {
                plan->graph = g;
                const hipError_t e3 = hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0);
                if (e3 != hipSuccess) rc = hip_err(e3, "hipGraphInstantiate");
            } else if (g) {
                (void)hipGraphDestroy(g);
            }
        } else {
            rc = hip_err(e, "hipStreamBeginCapture");
        }
        (void)hipStreamDestroy(cap);
        if (rc) return rc;
    }
    PLAN_HIP(hipGraphLaunch(plan->exec, static_cast<hipStream_t>(stream)), "hipGraphLaunch");
    return CRT_OK;
}

int64_t crt_batch_plan_rays(const crt_batch_plan* plan) { return plan ? plan->rays : 0; }

int crt_autotune_batched(const crt_batch* batches, int nbatches, const crt_batch_config* candidates,
                         int ncandidates, int repeats, void* stream, crt_batch_config* best, float* best_ms) {
    if (!best || !best_ms) return crt::set_error(CRT_E_INVALID, "crt_autotune_batched: null output");
    if (repeats < 1) repeats = 1;
    std::vector<crt_batch_config> cands;
    if (candidates && ncandidates > 0) {
        cands.assign(candidates, candidates + ncandidates);
    } else {
        // child order matters only to any-hit walks of the wide formats
        bool orders = false;
        for (int i = 0; i < nbatches; ++i)
            orders |= batches[i].query == CRT_QUERY_ANY &&
                      (batches[i].scene.format == CRT_FORMAT_WIDE4 || batches[i].scene.format == CRT_FORMAT_WIDE8);
        for (int v = 0; v < 12; ++v)
            for (int t : {1, 2, 4})
                for (int w : {1, 4})
                    for (int o : {0, 1})
                        for (int q : {1, crt::kQueueShards})
                            for (int co = 0; co <= (orders ? 1 : 0); ++co) {
                                crt_batch_config c{};
                                c.variant = v;
                                c.tiles_per_claim = t;
                                c.waves_per_group = w;
                                c.order = o;
                                c.shards = q;
                                c.child_order = co;
                                cands.push_back(c);
                            }
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PLAN_HIP(hipEventCreate(&e0), "hipEventCreate");
    PLAN_HIP(hipEventCreate(&e1), "hipEventCreate");
    float best_t = -1.0f;
    int rc = CRT_OK;
    for (const crt_batch_config& c : cands) {
        crt_batch_plan* p = nullptr;
        rc = crt_batch_plan_create(batches, nbatches, &c, &p);
        if (rc) break;
        rc = launch_plan(*p, s);  // warm-up
        if (!rc && hipEventRecord(e0, s) != hipSuccess) rc = crt::set_error(CRT_E_HIP, "hipEventRecord");
        for (int r = 0; !rc && r < repeats; ++r) rc = launch_plan(*p, s);
        if (!rc && hipEventRecord(e1, s) != hipSuccess) rc = crt::set_error(CRT_E_HIP, "hipEventRecord");
        float ms = 0.0f;
        if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = crt::set_error(CRT_E_HIP, "hipEventElapsedTime");
        delete p;
        if (rc) break;
        ms /= float(repeats);
        if (best_t < 0.0f || ms < best_t) {
            best_t = ms;
            *best = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (best_t < 0.0f)
        return crt::set_error(CRT_E_INVALID, "crt_autotune_batched: no candidate applies to these batches");
    *best_ms = best_t;
    return CRT_OK;
}

int crt_autotune_batched_layouts(const crt_batch* batches, int nbatches, int nlayouts,
                                 const crt_batch_config* candidates, int ncandidates, int repeats, void* stream,
                                 int32_t* best_layout, crt_batch_config* best, float* best_ms) {
    if (!batches || !best_layout || !best || !best_ms)
        return crt::set_error(CRT_E_INVALID, "crt_autotune_batched_layouts: null argument");
    if (nbatches < 1 || nlayouts < 1)
        return crt::set_error(CRT_E_INVALID, "crt_autotune_batched_layouts: needs batches and layouts");
    // every layout must hold the same work: same queries, flags, rays, hits and counts, batch by batch
    for (int l = 1; l < nlayouts; ++l)
        for (int i = 0; i < nbatches; ++i) {
            const crt_batch& a = batches[i];
            const crt_batch& b = batches[int64_t(l) * nbatches + i];
            if (a.query != b.query || a.flags != b.flags || a.rays != b.rays || a.hits != b.hits ||
                a.count != b.count || a.tri_tests != b.tri_tests)
                return crt::set_error(CRT_E_INVALID, "crt_autotune_batched_layouts: layout " + std::to_string(l) +
                                                         " batch " + std::to_string(i) +
                                                         " differs from layout 0 in more than its scene");
        }
    float best_t = -1.0f;
    for (int l = 0; l < nlayouts; ++l) {
        crt_batch_config c{};
        float ms = 0.0f;
        const int rc = crt_autotune_batched(batches + int64_t(l) * nbatches, nbatches, candidates, ncandidates,
                                            repeats, stream, &c, &ms);
        if (rc) return rc;
        if (best_t < 0.0f || ms < best_t) {
            best_t = ms;
            *best = c;
            *best_layout = l;
        }
    }
    *best_ms = best_t;
    return CRT_OK;
}

}  // extern "C"
