// Kernel arguments and launchers shared by the host runtime and trace.hip.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "candela_rt.h"

namespace crt {

// Device scene passed by value to every kernel (SSBO bindings 16..19 plus
// u_EntityCount / u_TotalNodes of the reference's traversal include).
struct SceneArgs {
    int32_t format;
    int32_t entity_count;
    int64_t total_nodes;
    const void* nodes;
    const crt_triangle* tris;
    const crt_vertex* verts;
    const crt_entity* entities;
    const crt_tri_record* tri_records;  // wide formats only
};

// SceneArgs of a scene view.
inline SceneArgs scene_args(const crt_scene_view& v) {
    return SceneArgs{v.format, v.entity_count, v.total_nodes, v.nodes, v.triangles, v.vertices, v.entities,
                     v.tri_records};
}

struct Camera {
    float inv_view[16];
    float inv_proj[16];
    float tmax;
};

constexpr int32_t kMaxIterations = 1024;  // TraverseBVHStackless.glsl:184

// The wave-scheduled walks (wave_walk.h, stack_walk.h, wide_walk.h,
// spec_walk.h) run the walk-end code (the next entity's transform, or the
// ray's hit record) once this many lanes have finished their walks, or once
// no lane can step: that code costs the whole wave however few lanes it
// serves.  Finished lanes only wait; a ray's record does not depend on when
// its end runs (the speculative twin replays the rule for its schedule).
// Measured on config 2: 4-8% off the sequential BVH4 walks, 3% off the BVH2
// wave walks, 6% off the speculative GI launch (16 and 48 gain less).
constexpr int32_t kWalkEndBatch = 32;

hipError_t launch_trace(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                        uint32_t flags, int32_t max_iter, int block, int lds_depth, uint32_t* tri_tests,
                        hipStream_t stream);

hipError_t launch_primary(const SceneArgs& s, const Camera& cam, crt_ray* rays, crt_hit* hits, int width, int height,
                          uint32_t flags, int32_t max_iter, int block, hipStream_t stream);

// Tuning of the persistent kernel.
struct PersistentShape {
    int variant;  // refill / inner / leaf-batch preset
    int grid;     // workgroups launched
};

hipError_t launch_persistent(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                             uint32_t flags, int32_t max_iter, unsigned long long* counter, uint32_t* tri_tests,
                             const PersistentShape& shape, hipStream_t stream);

hipError_t launch_grid(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n, uint32_t flags,
                       int32_t max_iter, uint32_t* tri_tests, int variant, hipStream_t stream);

// Refilling persistent walk of the wide layouts (variant 0..5; grid 0 = fill
// every CU); `counters` is a queue block (queue_acquire) the launch resets itself.
hipError_t launch_wide_persistent(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                                  uint32_t flags, int32_t max_iter, unsigned long long* counters, uint32_t* tri_tests,
                                  int variant, int grid, hipStream_t stream);

// Wave-scheduled walk of the wide layouts (variant 0..5).
hipError_t launch_wide_grid(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                            uint32_t flags, int32_t max_iter, uint32_t* tri_tests, int variant, hipStream_t stream);

// Wave-scheduled walk of the stack layout (variant 0..7).
hipError_t launch_stack_grid(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                             uint32_t flags, int32_t max_iter, uint32_t* tri_tests, int variant, hipStream_t stream);

// Refilling persistent walk of the stack layout (variant 0..7; grid 0 = fill
// every CU); `counters` is a queue block the launch resets itself.
hipError_t launch_stack_persistent(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                                   uint32_t flags, int32_t max_iter, unsigned long long* counters,
                                   uint32_t* tri_tests, int variant, int grid, hipStream_t stream);

// Speculative wave walk of the wide layouts (spec.hip; variant 0..5): grid
// launch, and device-filling waves that refill from a queue block.
hipError_t launch_spec_grid(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                            uint32_t flags, int32_t max_iter, uint32_t* tri_tests, int variant, hipStream_t stream);
hipError_t launch_spec_persistent(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                                  uint32_t flags, int32_t max_iter, unsigned long long* counters,
                                  uint32_t* tri_tests, int variant, int grid, hipStream_t stream);

// Speculative walks of the BVH2 layouts (spec_bvh2.hip; variant 0..3):
// stack = true for CRT_FORMAT_STACK, false for CRT_FORMAT_STACKLESS.
hipError_t launch_spec_bvh2_grid(const SceneArgs& s, bool stack, int query, const crt_ray* rays, crt_hit* hits,
                                 int64_t n, uint32_t flags, int32_t max_iter, uint32_t* tri_tests, int variant,
                                 hipStream_t stream);
hipError_t launch_spec_bvh2_persistent(const SceneArgs& s, bool stack, int query, const crt_ray* rays, crt_hit* hits,
                                       int64_t n, uint32_t flags, int32_t max_iter, unsigned long long* counters,
                                       uint32_t* tri_tests, int variant, int grid, hipStream_t stream);

// Workgroups of 256 threads that fill every CU with the persistent kernel.
int persistent_fill_grid(int query);

hipError_t launch_copy16(void* dst, const void* src, int64_t n16, hipStream_t stream);

}  // namespace crt
