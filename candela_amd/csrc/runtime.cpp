// Host runtime of libcandela_rt: errors, device memory, the raw trace entry
// and the RayIntersector<T> context (Source/Core/BVH/Intersector.h:60-351).
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "candela_rt.h"
#include "crt_internal.h"
#include "trace_launch.h"

namespace crt {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

static int hip_fail(hipError_t e, const char* what) {
    return set_error(CRT_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define CRT_HIP(call, what)                      \
    do {                                         \
        hipError_t e_ = (call);                  \
        if (e_ != hipSuccess) return ::crt::hip_fail(e_, what); \
    } while (0)

static int validate_trace_config(const crt_trace_config* cfg, int& block, int& depth) {
    block = 256;
    depth = 16;
    if (cfg) {
        if (cfg->block_size != 0) block = cfg->block_size;
        if (cfg->lds_stack != 0) depth = cfg->lds_stack;
    }
    if (block < 64 || block > 1024 || block % 64 != 0)
        return set_error(CRT_E_INVALID, "trace config: block_size must be a multiple of 64 in [64, 1024]");
    if (depth != 8 && depth != 16 && depth != 32 && depth != 48 && depth != 64)
        return set_error(CRT_E_INVALID, "trace config: lds_stack must be one of 8, 16, 32, 48, 64");
    if (size_t(depth) * size_t(block) * 4 > 64 * 1024)
        return set_error(CRT_E_INVALID, "trace config: lds_stack * block_size * 4 exceeds 64 KiB of LDS");
    return CRT_OK;
}

static bool is_wide(int format) { return format == CRT_FORMAT_WIDE4 || format == CRT_FORMAT_WIDE8; }

static int validate_scene(const crt_scene_view* s) {
    if (!s) return set_error(CRT_E_INVALID, "scene is null");
    if (s->format != CRT_FORMAT_STACKLESS && s->format != CRT_FORMAT_STACK && !is_wide(s->format))
        return set_error(CRT_E_INVALID, "scene: unknown node format");
    if (s->entity_count < 0) return set_error(CRT_E_INVALID, "scene: negative entity count");
    if (s->entity_count > 0 && is_wide(s->format) &&
        (!s->nodes || !s->tri_records || !s->entities || s->total_nodes <= 0))
        return set_error(CRT_E_INVALID, "scene: null device buffer (wide layout)");
    if (s->entity_count > 0 && !is_wide(s->format) &&
        (!s->nodes || !s->triangles || !s->vertices || !s->entities || s->total_nodes <= 0))
        return set_error(CRT_E_INVALID, "scene: null device buffer");
    return CRT_OK;
}

// Work-queue counters of the persistent kernels (queue.h).  A launch either
// brings its own block (crt_trace_config.queue: a graph that captures the
// launch owns its counters) or leases one from a per-device pool.  A pool
// block goes back to the pool when the launch has been enqueued, with a HIP
// event recorded after it; it is leased again only once that event has
// completed, so launches in flight on any number of streams never share
// counters.  When every block is busy the pool grows by kQueueChunk blocks,
// up to kQueueMaxBlocks, and then waits for the block leased longest ago.
constexpr int kQueueChunk = 64;
constexpr int kQueueMaxBlocks = 4096;
static_assert(size_t(kQueueBlockWords) * 8 == CRT_QUEUE_BYTES, "queue block size is the header's");

namespace {
struct QueuePool {
    std::mutex m;
    std::vector<unsigned long long*> blocks;
    std::vector<hipEvent_t> events;
    std::vector<uint8_t> state;   // 0 never used, 1 leased (not yet recorded), 2 recorded
    size_t next = 0;
};
QueuePool g_pools[64];
}  // namespace

int queue_acquire(const crt_trace_config* cfg, void* stream, const char* who, QueueLease& out) {
    out = QueueLease{};
    if (cfg && cfg->queue) {
        if (reinterpret_cast<uintptr_t>(cfg->queue) % 256 != 0)
            return set_error(CRT_E_INVALID, std::string(who) + ": config->queue must be 256-byte aligned");
        out.ptr = static_cast<unsigned long long*>(cfg->queue);
        return CRT_OK;
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    CRT_HIP(hipStreamIsCapturing(static_cast<hipStream_t>(stream), &cs), "hipStreamIsCapturing");
    if (cs != hipStreamCaptureStatusNone)
        return set_error(CRT_E_INVALID, std::string(who) +
                                            ": a persistent launch under stream capture needs config->queue "
                                            "(CRT_QUEUE_BYTES of device memory that the graph owns)");
    int dev = 0;
    CRT_HIP(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= 64) return set_error(CRT_E_INVALID, "device index out of range");
    QueuePool& P = g_pools[dev];
    std::lock_guard<std::mutex> lock(P.m);
    const size_t n = P.blocks.size();
    for (size_t k = 0; k < n; ++k) {
        const size_t i = (P.next + k) % n;
        if (P.state[i] == 1) continue;
        if (P.state[i] == 2) {
            const hipError_t q = hipEventQuery(P.events[i]);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) return hip_fail(q, "hipEventQuery(queue block)");
        }
        P.state[i] = 1;
        P.next = i + 1;
        out = QueueLease{P.blocks[i], dev, int32_t(i)};
        return CRT_OK;
    }
    if (n < size_t(kQueueMaxBlocks)) {
        void* p = nullptr;
        CRT_HIP(hipMalloc(&p, size_t(kQueueChunk) * CRT_QUEUE_BYTES), "hipMalloc(queue blocks)");
        for (int j = 0; j < kQueueChunk; ++j) {
            hipEvent_t e = nullptr;
            CRT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate(queue block)");
            P.blocks.push_back(reinterpret_cast<unsigned long long*>(static_cast<char*>(p) + size_t(j) * CRT_QUEUE_BYTES));
            P.events.push_back(e);
            P.state.push_back(0);
        }
        P.state[n] = 1;
        P.next = n + 1;
        out = QueueLease{P.blocks[n], dev, int32_t(n)};
        return CRT_OK;
    }
    // every block is in flight: wait for the next one in lease order
    const size_t i = P.next % n;
    CRT_HIP(hipEventSynchronize(P.events[i]), "hipEventSynchronize(queue block)");
    P.state[i] = 1;
    P.next = i + 1;
    out = QueueLease{P.blocks[i], dev, int32_t(i)};
    return CRT_OK;
}

int queue_release(const QueueLease& lease, void* stream) {
    if (lease.slot < 0) return CRT_OK;   // the caller's own block
    QueuePool& P = g_pools[lease.dev];
    std::lock_guard<std::mutex> lock(P.m);
    const hipError_t e = hipEventRecord(P.events[size_t(lease.slot)], static_cast<hipStream_t>(stream));
    // an unrecorded block stays leased (never reused) rather than be shared
    if (e != hipSuccess) return hip_fail(e, "hipEventRecord(queue block)");
    P.state[size_t(lease.slot)] = 2;
    return CRT_OK;
}

int trace(const crt_scene_view* scene, int query, uint32_t flags, const crt_ray* rays, crt_hit* hits, int64_t n,
          const crt_trace_config* cfg, void* stream) {
    int rc = validate_scene(scene);
    if (rc) return rc;
    if (query != CRT_QUERY_CLOSEST && query != CRT_QUERY_ANY) return set_error(CRT_E_INVALID, "crt_trace: bad query");
    if (n < 0) return set_error(CRT_E_INVALID, "crt_trace: negative ray count");
    if (n > 0 && (!rays || !hits)) return set_error(CRT_E_INVALID, "crt_trace: null ray or hit buffer");
    if (n > (int64_t(1) << 31) * 256) return set_error(CRT_E_INVALID, "crt_trace: too many rays for one launch");
    int block, depth;
    rc = validate_trace_config(cfg, block, depth);
    if (rc) return rc;
    const int kernel = cfg ? cfg->kernel : CRT_KERNEL_AUTO;
    if (kernel < CRT_KERNEL_AUTO || kernel > CRT_KERNEL_SPECULATIVE_REFILL)
        return set_error(CRT_E_INVALID, "crt_trace: unknown kernel");
    const bool speculative = kernel == CRT_KERNEL_SPECULATIVE || kernel == CRT_KERNEL_SPECULATIVE_REFILL;
    const bool bvh2 = scene->format == CRT_FORMAT_STACKLESS || scene->format == CRT_FORMAT_STACK;
    if (speculative && cfg && is_wide(scene->format) && cfg->variant > 5)
        return set_error(CRT_E_INVALID, "crt_trace: the speculative wide kernels take variant 0..5");
    if (speculative && cfg && bvh2 && cfg->variant > 3)
        return set_error(CRT_E_INVALID, "crt_trace: the speculative stack and stackless kernels take variant 0..3");
    if ((kernel == CRT_KERNEL_PERSISTENT || kernel == CRT_KERNEL_WAVE) && scene->format == CRT_FORMAT_STACK && cfg &&
        cfg->variant > 7)
        return set_error(CRT_E_INVALID, "crt_trace: the stack-layout wave kernels take variant 0..7");
    if (is_wide(scene->format) && (depth != 8 && depth != 16 && depth != 32))
        return set_error(CRT_E_INVALID, "crt_trace: wide formats take lds_stack 8, 16 or 32");
    if (is_wide(scene->format) && size_t(2 * depth + 8) * size_t(block) * 4 > 64 * 1024)
        return set_error(CRT_E_INVALID,
                         "trace config: (2 * lds_stack + width) * block_size * 4 exceeds 64 KiB of LDS");
    if (cfg && (cfg->variant < 0 || cfg->variant > 9 || cfg->grid < 0))
        return set_error(CRT_E_INVALID, "crt_trace: variant must be 0..9 and grid >= 0");
    const SceneArgs a = scene_args(*scene);
    uint32_t* tri_tests = cfg ? cfg->tri_tests : nullptr;
    const bool persistent = kernel == CRT_KERNEL_PERSISTENT;
    const bool wave = kernel == CRT_KERNEL_WAVE ||
                      (kernel == CRT_KERNEL_AUTO && (scene->format == CRT_FORMAT_STACKLESS || is_wide(scene->format)));
    if (n == 0) return CRT_OK;
    if (kernel == CRT_KERNEL_SPECULATIVE) {
        if (n >= (int64_t(1) << 31) * 256 - 2048)
            return set_error(CRT_E_INVALID, "crt_trace: too many rays for one launch");
        const int v = cfg ? cfg->variant : 0;
        hipStream_t st = static_cast<hipStream_t>(stream);
        CRT_HIP(bvh2 ? launch_spec_bvh2_grid(a, scene->format == CRT_FORMAT_STACK, query, rays, hits, n, flags,
                                             kMaxIterations, tri_tests, v, st)
                     : launch_spec_grid(a, query, rays, hits, n, flags, kMaxIterations, tri_tests, v, st),
                "crt_trace: speculative kernel launch");
        return CRT_OK;
    }
    if (kernel == CRT_KERNEL_SPECULATIVE_REFILL) {
        if (n >= (int64_t(1) << 31))
            return set_error(CRT_E_INVALID, "crt_trace: the persistent kernel takes < 2^31 rays per launch");
        QueueLease q;
        rc = queue_acquire(cfg, stream, "crt_trace", q);
        if (rc) return rc;
        const int v = cfg ? cfg->variant : 0, g = cfg ? cfg->grid : 0;
        hipStream_t st = static_cast<hipStream_t>(stream);
        const hipError_t e =
            bvh2 ? launch_spec_bvh2_persistent(a, scene->format == CRT_FORMAT_STACK, query, rays, hits, n, flags,
                                               kMaxIterations, q.ptr, tri_tests, v, g, st)
                 : launch_spec_persistent(a, query, rays, hits, n, flags, kMaxIterations, q.ptr, tri_tests, v, g, st);
        rc = queue_release(q, stream);
        CRT_HIP(e, "crt_trace: speculative persistent kernel launch");
        return rc;
    }
    if (wave && scene->format == CRT_FORMAT_STACK) {
        if (n >= (int64_t(1) << 31) * 256 - 2048)
            return set_error(CRT_E_INVALID, "crt_trace: too many rays for one launch");
        CRT_HIP(launch_stack_grid(a, query, rays, hits, n, flags, kMaxIterations, tri_tests, cfg ? cfg->variant : 0,
                                  static_cast<hipStream_t>(stream)),
                "crt_trace: stack wave kernel launch");
        return CRT_OK;
    }
    if (persistent && scene->format == CRT_FORMAT_STACK) {
        if (n >= (int64_t(1) << 31))
            return set_error(CRT_E_INVALID, "crt_trace: the persistent kernel takes < 2^31 rays per launch");
        QueueLease q;
        rc = queue_acquire(cfg, stream, "crt_trace", q);
        if (rc) return rc;
        const hipError_t e = launch_stack_persistent(a, query, rays, hits, n, flags, kMaxIterations, q.ptr, tri_tests,
                                                     cfg ? cfg->variant : 0, cfg ? cfg->grid : 0,
                                                     static_cast<hipStream_t>(stream));
        rc = queue_release(q, stream);
        CRT_HIP(e, "crt_trace: stack persistent kernel launch");
        return rc;
    }
    if (wave && is_wide(scene->format)) {
        if (n >= (int64_t(1) << 31) * 256 - 2048)
            return set_error(CRT_E_INVALID, "crt_trace: too many rays for one launch");
        CRT_HIP(launch_wide_grid(a, query, rays, hits, n, flags, kMaxIterations, tri_tests, cfg ? cfg->variant : 0,
                                 static_cast<hipStream_t>(stream)),
                "crt_trace: wide wave kernel launch");
        return CRT_OK;
    }
    if (wave) {
        if (n >= (int64_t(1) << 31) * 256 - 2048)
            return set_error(CRT_E_INVALID, "crt_trace: too many rays for one launch");
        CRT_HIP(launch_grid(a, query, rays, hits, n, flags, kMaxIterations, tri_tests, cfg ? cfg->variant : 0,
                            static_cast<hipStream_t>(stream)),
                "crt_trace: wave kernel launch");
        return CRT_OK;
    }
    if (persistent && is_wide(scene->format)) {
        if (n >= (int64_t(1) << 31))
            return set_error(CRT_E_INVALID, "crt_trace: the persistent kernel takes < 2^31 rays per launch");
        QueueLease q;
        rc = queue_acquire(cfg, stream, "crt_trace", q);
        if (rc) return rc;
        const hipError_t e = launch_wide_persistent(a, query, rays, hits, n, flags, kMaxIterations, q.ptr, tri_tests,
                                                    cfg ? cfg->variant : 0, cfg ? cfg->grid : 0,
                                                    static_cast<hipStream_t>(stream));
        rc = queue_release(q, stream);
        CRT_HIP(e, "crt_trace: wide persistent kernel launch");
        return rc;
    }
    if (persistent) {
        if (n >= (int64_t(1) << 31))
            return set_error(CRT_E_INVALID, "crt_trace: the persistent kernel takes < 2^31 rays per launch");
        QueueLease q;
        rc = queue_acquire(cfg, stream, "crt_trace", q);
        if (rc) return rc;
        PersistentShape shape{cfg ? cfg->variant : 0, cfg && cfg->grid ? cfg->grid : persistent_fill_grid(query)};
        const hipError_t e = launch_persistent(a, query, rays, hits, n, flags, kMaxIterations, q.ptr, tri_tests,
                                               shape, static_cast<hipStream_t>(stream));
        rc = queue_release(q, stream);
        CRT_HIP(e, "crt_trace: persistent kernel launch");
        return rc;
    }
    CRT_HIP(launch_trace(a, query, rays, hits, n, flags, kMaxIterations, block, depth, tri_tests,
                         static_cast<hipStream_t>(stream)),
            "crt_trace: kernel launch");
    return CRT_OK;
}

}  // namespace crt

// ---------------------------------------------------------------------------
// RayIntersector<T> context
// ---------------------------------------------------------------------------
struct ObjectData {  // _ObjectData, Intersector.h:51-56
    int64_t triangle_offset;
    int64_t vertices_offset;
    int64_t node_offset;
    int64_t node_count;
};

struct crt_context {
    int format = CRT_FORMAT_STACKLESS;
    // Host copies (m_BVHNodes, m_BVHVertices, m_BVHTriangles, m_BVHEntities).
    std::vector<crt_stackless_node> stackless;
    std::vector<crt_stack_node> stack;
    std::vector<crt_vertex> vertices;
    std::vector<crt_triangle> triangles;
    std::vector<crt_entity> entities;         // buffered (m_BVHEntities)
    std::vector<crt_entity> pending;          // m_Entities
    std::map<int32_t, ObjectData> objects;    // m_ObjectData
    uint32_t index_offset = 0;                // m_IndexOffset
    // Device buffers.
    void* d_nodes = nullptr;
    void* d_tris = nullptr;
    void* d_verts = nullptr;
    void* d_entities = nullptr;
    int64_t nodes_buffered = 0;               // m_NodeCountBuffered
    int64_t entities_buffered = 0;            // m_EntityPushed
    bool data_buffered = false;
    // Wide layout derived from the buffered data and entities.
    void* d_wide_groups = nullptr;
    void* d_wide_records = nullptr;
    void* d_wide_entities = nullptr;
    int32_t wide_width = 0;                   // 0 = not built or stale
    int64_t wide_nodes = 0;
    int64_t wide_entities = 0;

    int64_t node_count() const {
        return format == CRT_FORMAT_STACKLESS ? int64_t(stackless.size()) : int64_t(stack.size());
    }
    void free_device() {
        if (d_nodes) (void)hipFree(d_nodes);
        if (d_tris) (void)hipFree(d_tris);
        if (d_verts) (void)hipFree(d_verts);
        if (d_entities) (void)hipFree(d_entities);
        d_nodes = d_tris = d_verts = d_entities = nullptr;
        free_wide();
    }
    void free_wide() {
        if (d_wide_groups) (void)hipFree(d_wide_groups);
        if (d_wide_records) (void)hipFree(d_wide_records);
        if (d_wide_entities) (void)hipFree(d_wide_entities);
        d_wide_groups = d_wide_records = d_wide_entities = nullptr;
        wide_width = 0;
    }
};

namespace {

// glm::inverse of a float mat4, as PushEntity stores it (Intersector.h:209):
// the steps of glm 0.9.8.5's generic compute_inverse<tmat4x4>
// (glm/detail/func_matrix.inl:297-351) in float, in the same order, so the
// inverse is byte-identical to the reference's.  The 2x2 minors of the lower
// rows (coef), the cofactor columns (inv_k = v1*f0 - v2*f1 + v3*f2 and the
// like, signs alternating), the determinant as the first column of the input
// against the first row of the adjugate summed pairwise, then one reciprocal
// multiplied into every entry.  `a` and `out` are column-major (a[c * 4 + r]).
// The reference pushes the inf / NaN entries of a singular matrix; this
// returns false instead.
bool invert4(const float* a, float* out) {
    auto m = [a](int c, int r) { return a[c * 4 + r]; };
    const float c00 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3);
    const float c02 = m(1, 2) * m(3, 3) - m(3, 2) * m(1, 3);
    const float c03 = m(1, 2) * m(2, 3) - m(2, 2) * m(1, 3);
    const float c04 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3);
    const float c06 = m(1, 1) * m(3, 3) - m(3, 1) * m(1, 3);
    const float c07 = m(1, 1) * m(2, 3) - m(2, 1) * m(1, 3);
    const float c08 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2);
    const float c10 = m(1, 1) * m(3, 2) - m(3, 1) * m(1, 2);
    const float c11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2);
    const float c12 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3);
    const float c14 = m(1, 0) * m(3, 3) - m(3, 0) * m(1, 3);
    const float c15 = m(1, 0) * m(2, 3) - m(2, 0) * m(1, 3);
    const float c16 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2);
    const float c18 = m(1, 0) * m(3, 2) - m(3, 0) * m(1, 2);
    const float c19 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2);
    const float c20 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1);
    const float c22 = m(1, 0) * m(3, 1) - m(3, 0) * m(1, 1);
    const float c23 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
    const float f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07}, f2[4] = {c08, c08, c10, c11};
    const float f3[4] = {c12, c12, c14, c15}, f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
    float v[4][4];   // v[r] = (m[1][r], m[0][r], m[0][r], m[0][r])
    for (int r = 0; r < 4; ++r) {
        v[r][0] = m(1, r);
        v[r][1] = v[r][2] = v[r][3] = m(0, r);
    }
    float inv[4][4];   // adjugate columns
    for (int i = 0; i < 4; ++i) {
        const float sa = (i % 2 == 0) ? 1.0f : -1.0f;   // (+1, -1, +1, -1)
        const float sb = -sa;                             // (-1, +1, -1, +1)
        inv[0][i] = (v[1][i] * f0[i] - v[2][i] * f1[i] + v[3][i] * f2[i]) * sa;
        inv[1][i] = (v[0][i] * f0[i] - v[2][i] * f3[i] + v[3][i] * f4[i]) * sb;
        inv[2][i] = (v[0][i] * f1[i] - v[1][i] * f3[i] + v[3][i] * f5[i]) * sa;
        inv[3][i] = (v[0][i] * f2[i] - v[1][i] * f4[i] + v[2][i] * f5[i]) * sb;
    }
    float d[4];
    for (int c = 0; c < 4; ++c) d[c] = m(0, c) * inv[c][0];
    const float det = (d[0] + d[1]) + (d[2] + d[3]);
    if (det == 0.0f || !std::isfinite(det)) return false;
    const float one = 1.0f / det;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) out[c * 4 + r] = inv[c][r] * one;
    return true;
}

template <typename T>
int upload(void*& dst, const std::vector<T>& src) {
    if (dst) {
        (void)hipFree(dst);
        dst = nullptr;
    }
    const size_t bytes = std::max<size_t>(sizeof(T), src.size() * sizeof(T));
    CRT_HIP(hipMalloc(&dst, bytes), "hipMalloc");
    if (!src.empty()) CRT_HIP(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
    return CRT_OK;
}

// Appends one object's buffers with the offset rules of AddObject
// (Intersector.h:178-197).
int append_object(crt_context* ctx, int32_t object_id, const void* nodes, int64_t node_count,
                  const crt_triangle* tris, int64_t tri_count, const crt_vertex* verts, int64_t vert_count) {
    ObjectData od;
    od.node_offset = ctx->node_count();
    od.triangle_offset = int64_t(ctx->triangles.size());
    od.vertices_offset = int64_t(ctx->vertices.size());
    od.node_count = node_count;
    if (ctx->format == CRT_FORMAT_STACKLESS) {
        const auto* n = static_cast<const crt_stackless_node*>(nodes);
        ctx->stackless.insert(ctx->stackless.end(), n, n + node_count);
    } else {
        const auto* n = static_cast<const crt_stack_node*>(nodes);
        ctx->stack.insert(ctx->stack.end(), n, n + node_count);
    }
    ctx->vertices.insert(ctx->vertices.end(), verts, verts + vert_count);
    ctx->triangles.reserve(ctx->triangles.size() + size_t(tri_count));
    for (int64_t i = 0; i < tri_count; ++i) {
        crt_triangle t = tris[i];
        t.packed[0] += int32_t(ctx->index_offset);
        t.packed[1] += int32_t(ctx->index_offset);
        t.packed[2] += int32_t(ctx->index_offset);
        ctx->triangles.push_back(t);
    }
    ctx->index_offset += uint32_t(vert_count);
    ctx->objects[object_id] = od;
    ctx->data_buffered = false;
    return CRT_OK;
}

}  // namespace

extern "C" {

const char* crt_last_error(void) { return crt::g_last_error.c_str(); }
int crt_abi_version(void) { return CRT_ABI_VERSION; }

int crt_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int crt_device_alloc(void** ptr, size_t bytes) {
    if (!ptr) return crt::set_error(CRT_E_INVALID, "crt_device_alloc: ptr is null");
    *ptr = nullptr;
    if (bytes == 0) bytes = 1;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return crt::set_error(CRT_E_NOMEM, "crt_device_alloc: out of device memory");
    if (e != hipSuccess) return crt::hip_fail(e, "crt_device_alloc");
    return CRT_OK;
}

int crt_device_free(void* ptr) {
    if (!ptr) return CRT_OK;
    CRT_HIP(hipFree(ptr), "crt_device_free");
    return CRT_OK;
}

int crt_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return CRT_OK;
    if (!dst || !src) return crt::set_error(CRT_E_INVALID, "crt_memcpy_h2d: null pointer");
    CRT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)), "hipMemcpyAsync");
    CRT_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize");
    return CRT_OK;
}

int crt_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return CRT_OK;
    if (!dst || !src) return crt::set_error(CRT_E_INVALID, "crt_memcpy_d2h: null pointer");
    CRT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream)), "hipMemcpyAsync");
    CRT_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize");
    return CRT_OK;
}

int crt_memset_d(void* dst, int value, size_t bytes, void* stream) {
    if (bytes == 0) return CRT_OK;
    if (!dst) return crt::set_error(CRT_E_INVALID, "crt_memset_d: null pointer");
    CRT_HIP(hipMemsetAsync(dst, value, bytes, static_cast<hipStream_t>(stream)), "hipMemsetAsync");
    return CRT_OK;
}

int crt_device_copy(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return CRT_OK;
    if (!dst || !src || bytes % 16 || (reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) % 16)
        return crt::set_error(CRT_E_INVALID, "crt_device_copy: needs 16-byte aligned pointers and size");
    CRT_HIP(crt::launch_copy16(dst, src, int64_t(bytes / 16), static_cast<hipStream_t>(stream)),
            "crt_device_copy: kernel launch");
    return CRT_OK;
}

int crt_stream_create(void** stream) {
    if (!stream) return crt::set_error(CRT_E_INVALID, "crt_stream_create: null out");
    hipStream_t s = nullptr;
    CRT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    *stream = s;
    return CRT_OK;
}

int crt_stream_destroy(void* stream) {
    if (!stream) return CRT_OK;
    CRT_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)), "hipStreamDestroy");
    return CRT_OK;
}

int crt_stream_sync(void* stream) {
    CRT_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize");
    return CRT_OK;
}

int crt_device_sync(void) {
    CRT_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return CRT_OK;
}

int crt_graph_begin(void* stream) {
    if (!stream) return crt::set_error(CRT_E_INVALID, "crt_graph_begin: capture needs a created stream, not the null stream");
    CRT_HIP(hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal),
            "hipStreamBeginCapture");
    return CRT_OK;
}

int crt_graph_end(void* stream, void** graph) {
    if (!stream || !graph) return crt::set_error(CRT_E_INVALID, "crt_graph_end: null argument");
    *graph = nullptr;
    hipGraph_t g = nullptr;
    CRT_HIP(hipStreamEndCapture(static_cast<hipStream_t>(stream), &g), "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    CRT_HIP(e, "hipGraphInstantiate");
    *graph = exec;
    return CRT_OK;
}

int crt_graph_launch(void* graph, void* stream) {
    if (!graph) return crt::set_error(CRT_E_INVALID, "crt_graph_launch: null graph");
    CRT_HIP(hipGraphLaunch(static_cast<hipGraphExec_t>(graph), static_cast<hipStream_t>(stream)), "hipGraphLaunch");
    return CRT_OK;
}

int crt_graph_destroy(void* graph) {
    if (!graph) return CRT_OK;
    CRT_HIP(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph)), "hipGraphExecDestroy");
    return CRT_OK;
}

int crt_event_create(void** event) {
    if (!event) return crt::set_error(CRT_E_INVALID, "crt_event_create: null out");
    hipEvent_t e = nullptr;
    CRT_HIP(hipEventCreate(&e), "hipEventCreate");
    *event = e;
    return CRT_OK;
}

int crt_event_destroy(void* event) {
    if (!event) return CRT_OK;
    CRT_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)), "hipEventDestroy");
    return CRT_OK;
}

int crt_event_record(void* event, void* stream) {
    if (!event) return crt::set_error(CRT_E_INVALID, "crt_event_record: null event");
    CRT_HIP(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)), "hipEventRecord");
    return CRT_OK;
}

int crt_stream_wait_event(void* stream, void* event) {
    if (!event) return crt::set_error(CRT_E_INVALID, "crt_stream_wait_event: null event");
    CRT_HIP(hipStreamWaitEvent(static_cast<hipStream_t>(stream), static_cast<hipEvent_t>(event), 0),
            "hipStreamWaitEvent");
    return CRT_OK;
}

int crt_event_elapsed_ms(void* start, void* end, float* ms) {
    if (!start || !end || !ms) return crt::set_error(CRT_E_INVALID, "crt_event_elapsed_ms: null argument");
    CRT_HIP(hipEventSynchronize(static_cast<hipEvent_t>(end)), "hipEventSynchronize");
    CRT_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(end)),
            "hipEventElapsedTime");
    return CRT_OK;
}

int crt_trace(const crt_scene_view* scene, int query, uint32_t flags, const crt_ray* rays, crt_hit* hits,
              int64_t nrays, const crt_trace_config* config, void* stream) {
    return crt::trace(scene, query, flags, rays, hits, nrays, config, stream);
}

// --- context ----------------------------------------------------------------

int crt_create(int format, crt_context** out) {
    if (!out) return crt::set_error(CRT_E_INVALID, "crt_create: out is null");
    *out = nullptr;
    if (format != CRT_FORMAT_STACKLESS && format != CRT_FORMAT_STACK)
        return crt::set_error(CRT_E_INVALID, "crt_create: unknown node format");
    crt_context* c = new (std::nothrow) crt_context();
    if (!c) return crt::set_error(CRT_E_NOMEM, "crt_create: out of host memory");
    c->format = format;
    *out = c;
    return CRT_OK;
}

void crt_destroy(crt_context* ctx) {
    if (!ctx) return;
    ctx->free_device();
    delete ctx;
}

int crt_context_format(const crt_context* ctx) { return ctx ? ctx->format : CRT_E_INVALID; }

int crt_add_object(crt_context* ctx, int32_t object_id, const crt_mesh* meshes, int mesh_count,
                   const crt_build_options* options) {
    if (!ctx) return crt::set_error(CRT_E_INVALID, "crt_add_object: ctx is null");
    crt::BvhResult r;
    // t_offset = triangles already added (Intersector.h:183).
    int rc = crt::build_bvh(ctx->format, meshes, mesh_count, int32_t(ctx->triangles.size()), options, r);
    if (rc) return rc;
    const void* nodes = ctx->format == CRT_FORMAT_STACKLESS ? static_cast<const void*>(r.stackless.data())
                                                            : static_cast<const void*>(r.stack.data());
    return append_object(ctx, object_id, nodes, int64_t(r.stats.array_length), r.triangles.data(),
                         int64_t(r.triangles.size()), r.vertices.data(), int64_t(r.vertices.size()));
}

int crt_add_object_prebuilt(crt_context* ctx, int32_t object_id, const void* nodes, int64_t node_count,
                            const crt_triangle* triangles, int64_t triangle_count, const crt_vertex* vertices,
                            int64_t vertex_count) {
    if (!ctx) return crt::set_error(CRT_E_INVALID, "crt_add_object_prebuilt: ctx is null");
    if (node_count <= 0 || !nodes || triangle_count < 0 || (triangle_count && !triangles) || vertex_count < 0 ||
        (vertex_count && !vertices))
        return crt::set_error(CRT_E_INVALID, "crt_add_object_prebuilt: bad buffers");
    return append_object(ctx, object_id, nodes, node_count, triangles, triangle_count, vertices, vertex_count);
}

int crt_add_object_cached(crt_context* ctx, int32_t object_id, const crt_mesh* meshes, int mesh_count,
                          const crt_build_options* options, const char* cache_dir, int32_t* cache_hit) {
    if (cache_hit) *cache_hit = 0;
    if (!ctx || !cache_dir) return crt::set_error(CRT_E_INVALID, "crt_add_object_cached: null argument");
    uint64_t key = 0;
    int rc = crt::bvh_key(ctx->format, meshes, mesh_count, options, &key);
    if (rc) return rc;
    char name[32];
    std::snprintf(name, sizeof(name), "%016llx.crtbvh", static_cast<unsigned long long>(key));
    std::string path(cache_dir);
    if (!path.empty() && path.back() != '/') path += '/';
    path += name;
    const int32_t t_offset = int32_t(ctx->triangles.size());
    const bool seeded = options && options->child_swap == CRT_SWAP_SEEDED;

    crt_bvh b;
    crt_cache_header hd;
    bool hit = crt::load_bvh(path, key, b, &hd) == CRT_OK && hd.format == ctx->format &&
               (!seeded || crt::same_toolchain(hd));
    int32_t status = 1;
    if (!hit) {
        b = crt_bvh();
        rc = crt::build_bvh(ctx->format, meshes, mesh_count, 0, options, b.r);
        if (rc) return rc;
        b.t_offset = 0;
        if (options) b.options = *options;
        status = crt::save_bvh(b, key, path) == CRT_OK ? 0 : -1;
    }
    std::string save_error = status < 0 ? crt_last_error() : std::string();
    rc = crt::rebase_bvh(b, t_offset);
    if (rc) return rc;
    const void* nodes = ctx->format == CRT_FORMAT_STACKLESS ? static_cast<const void*>(b.r.stackless.data())
                                                            : static_cast<const void*>(b.r.stack.data());
    rc = append_object(ctx, object_id, nodes, int64_t(b.r.stats.array_length), b.r.triangles.data(),
                       int64_t(b.r.triangles.size()), b.r.vertices.data(), int64_t(b.r.vertices.size()));
    if (rc) return rc;
    if (status < 0) crt::set_error(CRT_E_IO, save_error);
    if (cache_hit) *cache_hit = status;
    return CRT_OK;
}

int crt_push_entity(crt_context* ctx, int32_t object_id, const float model[16], float emissive, float translucency) {
    if (!ctx || !model) return crt::set_error(CRT_E_INVALID, "crt_push_entity: null argument");
    auto it = ctx->objects.find(object_id);
    if (it == ctx->objects.end())
        return crt::set_error(CRT_E_INVALID,
                              "Trying to push entity whose parent object hasn't been added to global BVH");
    crt_entity e;
    const int rc = crt_make_entity(model, emissive, translucency, int32_t(it->second.node_offset),
                                   int32_t(it->second.node_count), &e);
    if (rc) return rc;
    ctx->pending.push_back(e);
    return CRT_OK;
}

int crt_make_entity(const float model[16], float emissive, float translucency, int32_t node_offset,
                    int32_t node_count, crt_entity* out) {
    if (!model || !out) return crt::set_error(CRT_E_INVALID, "crt_make_entity: null argument");
    crt_entity e;
    std::memset(&e, 0, sizeof(e));
    std::memcpy(e.model, model, sizeof(e.model));
    if (!invert4(model, e.inverse)) return crt::set_error(CRT_E_INVALID, "crt_make_entity: singular model matrix");
    e.node_offset = node_offset;
    e.node_count = node_count;
    const float alpha = 1.0f - translucency;
    std::memcpy(&e.data[0], &emissive, 4);
    std::memcpy(&e.data[1], &alpha, 4);
    *out = e;
    return CRT_OK;
}

int crt_buffer_entities(crt_context* ctx) {
    if (!ctx) return crt::set_error(CRT_E_INVALID, "crt_buffer_entities: ctx is null");
    int rc = upload(ctx->d_entities, ctx->pending);
    if (rc) return rc;
    ctx->entities_buffered = int64_t(ctx->pending.size());
    ctx->entities = ctx->pending;
    ctx->pending.clear();
    ctx->free_wide();
    return CRT_OK;
}

int crt_buffer_data(crt_context* ctx, int clear_cpu_data) {
    if (!ctx) return crt::set_error(CRT_E_INVALID, "crt_buffer_data: ctx is null");
    int rc = upload(ctx->d_tris, ctx->triangles);
    if (!rc) rc = ctx->format == CRT_FORMAT_STACKLESS ? upload(ctx->d_nodes, ctx->stackless) : upload(ctx->d_nodes, ctx->stack);
    if (!rc) rc = upload(ctx->d_verts, ctx->vertices);
    if (rc) return rc;
    ctx->nodes_buffered = ctx->node_count();
    ctx->data_buffered = true;
    ctx->free_wide();
    if (clear_cpu_data) {
        ctx->stackless.clear();
        ctx->stack.clear();
        ctx->vertices.clear();
        ctx->triangles.clear();
    }
    return CRT_OK;
}

int crt_context_host_buffers(const crt_context* ctx, const void** nodes, int64_t* node_count,
                             const crt_triangle** tris, int64_t* tri_count, const crt_vertex** verts,
                             int64_t* vert_count, const crt_entity** entities, int64_t* entity_count) {
    if (!ctx) return crt::set_error(CRT_E_INVALID, "crt_context_host_buffers: ctx is null");
    if (nodes)
        *nodes = ctx->format == CRT_FORMAT_STACKLESS ? static_cast<const void*>(ctx->stackless.data())
                                                     : static_cast<const void*>(ctx->stack.data());
    if (node_count) *node_count = ctx->node_count();
    if (tris) *tris = ctx->triangles.data();
    if (tri_count) *tri_count = int64_t(ctx->triangles.size());
    if (verts) *verts = ctx->vertices.data();
    if (vert_count) *vert_count = int64_t(ctx->vertices.size());
    if (entities) *entities = ctx->entities.data();
    if (entity_count) *entity_count = int64_t(ctx->entities.size());
    return CRT_OK;
}

int crt_context_scene(const crt_context* ctx, crt_scene_view* out) {
    if (!ctx || !out) return crt::set_error(CRT_E_INVALID, "crt_context_scene: null argument");
    if (!ctx->data_buffered || !ctx->d_entities)
        return crt::set_error(CRT_E_STATE, "crt_context_scene: call crt_buffer_data and crt_buffer_entities first");
    out->format = ctx->format;
    out->entity_count = int32_t(ctx->entities_buffered);
    out->total_nodes = ctx->nodes_buffered;
    out->nodes = ctx->d_nodes;
    out->triangles = static_cast<const crt_triangle*>(ctx->d_tris);
    out->vertices = static_cast<const crt_vertex*>(ctx->d_verts);
    out->entities = static_cast<const crt_entity*>(ctx->d_entities);
    out->tri_records = nullptr;
    return CRT_OK;
}

int crt_context_build_wide(crt_context* ctx, int32_t width, int32_t max_leaf) {
    const crt_wide_options o{width, max_leaf, CRT_COLLAPSE_GREEDY, 0.0f, 0.0f};
    return crt_context_build_wide_ex(ctx, &o);
}

int crt_context_build_wide_ex(crt_context* ctx, const crt_wide_options* options) {
    if (!ctx) return crt::set_error(CRT_E_INVALID, "crt_context_build_wide: ctx is null");
    if (!options) return crt::set_error(CRT_E_INVALID, "crt_context_build_wide: options is null");
    const int32_t width = options->width;
    if (!ctx->data_buffered || !ctx->d_entities)
        return crt::set_error(CRT_E_STATE, "crt_context_build_wide: call crt_buffer_data and crt_buffer_entities first");
    if (ctx->node_count() != ctx->nodes_buffered)
        return crt::set_error(CRT_E_STATE,
                              "crt_context_build_wide: the host copies were cleared or changed since crt_buffer_data");
    const void* nodes = ctx->format == CRT_FORMAT_STACKLESS ? static_cast<const void*>(ctx->stackless.data())
                                                            : static_cast<const void*>(ctx->stack.data());
    crt_wide* w = nullptr;
    int rc = crt_wide_build_ex(ctx->format, nodes, ctx->node_count(), ctx->triangles.data(),
                               int64_t(ctx->triangles.size()), ctx->vertices.data(), int64_t(ctx->vertices.size()),
                               ctx->entities.data(), int32_t(ctx->entities.size()), options, &w);
    if (rc) return rc;
    crt_wide_stats st;
    crt_wide_get_stats(w, &st);
    ctx->free_wide();
    const crt_wide_group* g = crt_wide_groups(w);
    const crt_tri_record* r = crt_wide_records(w);
    const crt_entity* e = crt_wide_entities(w);
    std::vector<crt_wide_group> vg(g, g + st.node_count * (width / 4));
    std::vector<crt_tri_record> vr(r, r + st.record_count);
    std::vector<crt_entity> ve(e, e + ctx->entities.size());
    crt_wide_free(w);
    rc = upload(ctx->d_wide_groups, vg);
    if (!rc) rc = upload(ctx->d_wide_records, vr);
    if (!rc) rc = upload(ctx->d_wide_entities, ve);
    if (rc) {
        ctx->free_wide();
        return rc;
    }
    ctx->wide_width = width;
    ctx->wide_nodes = st.node_count;
    ctx->wide_entities = int64_t(ve.size());
    return CRT_OK;
}

int crt_context_wide_scene(const crt_context* ctx, crt_scene_view* out) {
    if (!ctx || !out) return crt::set_error(CRT_E_INVALID, "crt_context_wide_scene: null argument");
    if (!ctx->wide_width)
        return crt::set_error(CRT_E_STATE, "crt_context_wide_scene: call crt_context_build_wide first");
    out->format = ctx->wide_width == 8 ? CRT_FORMAT_WIDE8 : CRT_FORMAT_WIDE4;
    out->entity_count = int32_t(ctx->wide_entities);
    out->total_nodes = ctx->wide_nodes;
    out->nodes = ctx->d_wide_groups;
    out->triangles = static_cast<const crt_triangle*>(ctx->d_tris);
    out->vertices = static_cast<const crt_vertex*>(ctx->d_verts);
    out->entities = static_cast<const crt_entity*>(ctx->d_wide_entities);
    out->tri_records = static_cast<const crt_tri_record*>(ctx->d_wide_records);
    return CRT_OK;
}

int crt_intersect(crt_context* ctx, int query, uint32_t flags, const crt_ray* rays, crt_hit* hits, int64_t nrays,
                  void* stream) {
    crt_scene_view s;
    int rc = crt_context_scene(ctx, &s);
    if (rc) return rc;
    return crt::trace(&s, query, flags, rays, hits, nrays, nullptr, stream);
}

int crt_intersect_primary(crt_context* ctx, crt_ray* rays, crt_hit* hits, int width, int height,
                          const float inv_view[16], const float inv_proj[16], uint32_t flags, void* stream) {
    crt_scene_view s;
    int rc = crt_context_scene(ctx, &s);
    if (rc) return rc;
    if (width <= 0 || height <= 0 || !rays || !hits || !inv_view || !inv_proj)
        return crt::set_error(CRT_E_INVALID, "crt_intersect_primary: bad arguments");
    crt::Camera cam;
    std::memcpy(cam.inv_view, inv_view, sizeof(cam.inv_view));
    std::memcpy(cam.inv_proj, inv_proj, sizeof(cam.inv_proj));
    cam.tmax = 1000000.0f;  // IntersectScene's TMax (TraverseBVHStackless.glsl:309)
    const crt::SceneArgs a = crt::scene_args(s);
    CRT_HIP(crt::launch_primary(a, cam, rays, hits, width, height, flags, crt::kMaxIterations, 256,
                                static_cast<hipStream_t>(stream)),
            "crt_intersect_primary: kernel launch");
    return CRT_OK;
}

}  // extern "C"
