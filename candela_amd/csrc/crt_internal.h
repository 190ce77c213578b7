// Internal declarations shared by the host translation units of libcandela_rt.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "candela_rt.h"

struct crt_bvh;

namespace crt {

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string& msg);

struct BvhResult {
    int format = CRT_FORMAT_STACKLESS;
    std::vector<crt_stackless_node> stackless;
    std::vector<crt_stack_node> stack;
    std::vector<crt_triangle> triangles;
    std::vector<crt_vertex> vertices;
    crt_bvh_stats stats{};
};

int build_bvh(int format, const crt_mesh* meshes, int mesh_count, int32_t t_offset,
              const crt_build_options* options, BvhResult& res);

// Binned-SAH constants of the constructor (BVHConstructor.cpp:45,49).
constexpr int kBinCount = 64;
constexpr uint32_t kMaxLeafTris = 2;

// BVH cache files (scene_cache.cpp).
int bvh_key(int format, const crt_mesh* meshes, int mesh_count, const crt_build_options* options, uint64_t* key);
int save_bvh(const crt_bvh& b, uint64_t key, const std::string& path);
int load_bvh(const std::string& path, uint64_t expect_key, crt_bvh& b, crt_cache_header* header_out);
int rebase_bvh(crt_bvh& b, int32_t t_offset);
// True when the header's toolchain is this library's (CRT_SWAP_SEEDED draws match).
bool same_toolchain(const crt_cache_header& hd);

// Counters of one persistent launch (queue.h): kQueueBlockWords u64, zeroed
// by the launch that uses them.  queue_acquire takes config->queue when the
// caller brings one, else leases a block from the per-device pool (refused
// under stream capture); queue_release, called once the launch is enqueued,
// records the event that frees a leased block.
constexpr int kQueueBlockWords = 256;
struct QueueLease {
    unsigned long long* ptr = nullptr;
    int dev = 0;
    int32_t slot = -1;   // pool block, or -1 for the caller's own
};
int queue_acquire(const crt_trace_config* cfg, void* stream, const char* who, QueueLease& out);
int queue_release(const QueueLease& lease, void* stream);

// crt_trace's body (runtime.cpp); crt_trace_noalloc (noalloc.hip) calls it with the static queue block.
int trace(const crt_scene_view* scene, int query, uint32_t flags, const crt_ray* rays, crt_hit* hits, int64_t n,
          const crt_trace_config* cfg, void* stream);

}  // namespace crt

// A built object: its flattened buffers and what they were built with.
struct crt_bvh {
    crt::BvhResult r;
    int32_t t_offset = 0;
    crt_build_options options{CRT_SWAP_NONE, 0u, 0};
};
