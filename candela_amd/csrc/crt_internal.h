// Internal declarations shared by the host translation units of libcandela_rt.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "candela_rt.h"

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

// A zeroed-by-the-launch device counter from the per-device pool of the
// persistent kernels (one slot per launch, handed out round robin).
int counter_slot(unsigned long long** out);

}  // namespace crt
