// crt_trace_noalloc: crt_trace for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, hosts that forbid
// allocation after start-up).
//
// crt_trace's persistent kernels (CRT_KERNEL_PERSISTENT over every layout and
// CRT_KERNEL_SPECULATIVE_REFILL) take one block of work-queue counters
// (CRT_QUEUE_BYTES: 16 shard counters on their own 128-byte lines, queue.h).
// crt_trace leases it from a per-device pool, which allocates when the pool
// grows, or takes the caller's config->queue.  Here the block is `scratch`, a
// file-scope device array: each GPU the library runs on holds its own copy
// from module load until the process ends, so a call allocates nothing.  The
// kernels zero the counters themselves (a one-workgroup reset kernel on the
// call's stream, no memset), so the block may hold anything the last call
// left.  Every call on one GPU shares that GPU's block: calls whose persistent
// launches overlap on one GPU race on the counters and lose or repeat rays.
//
// In host code `scratch` names a host-side stand-in, not the device copy:
// its device address comes from hipGetSymbolAddress, taken once per device on
// the first call made with that device current, and only that address is
// passed to the kernels.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "candela_rt.h"
#include "crt_internal.h"

namespace crt {
namespace {

static_assert(CRT_QUEUE_BYTES % 256 == 0, "the queue block is a whole number of 256-byte lines");

// One launch's work-queue counters.
static __device__ __align__(256) unsigned long long scratch[CRT_QUEUE_BYTES / 8];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};   // device address of `scratch` on each device, once taken

// The current device's address of `scratch`.
int scratch_address(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return set_error(CRT_E_HIP, std::string("crt_trace_noalloc: hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return set_error(CRT_E_INVALID, "crt_trace_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(CRT_E_HIP,
                             std::string("crt_trace_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return CRT_OK;
}

}  // namespace
}  // namespace crt

extern "C" {

int crt_trace_noalloc(const crt_scene_view* scene, int query, uint32_t flags, const crt_ray* rays, crt_hit* hits,
                      int64_t nrays, const crt_trace_config* config, void* stream) {
    if (nrays > CRT_NOALLOC_MAX_RAYS)
        return crt::set_error(CRT_E_INVALID, "crt_trace_noalloc: more than CRT_NOALLOC_MAX_RAYS (2^31 - 1) rays; "
                                             "the static queue block indexes one launch of at most that many");
    if (config && config->queue)
        return crt::set_error(CRT_E_INVALID, "crt_trace_noalloc: config->queue must be NULL (the counters are the "
                                             "library's static block; use crt_trace to bring your own)");
    void* block = nullptr;
    const int rc = crt::scratch_address(&block);
    if (rc) return rc;
    crt_trace_config cfg{};
    if (config) cfg = *config;
    cfg.queue = block;
    return crt::trace(scene, query, flags, rays, hits, nrays, &cfg, stream);
}

}  // extern "C"
