// Persistent-wave traversal with lane refill and batched leaf tests (gfx950).
//
// The same per-ray computation as trace.hip -- IntersectBVHStackless /
// IntersectBVHStacklessOcclusion inside the IntersectScene entity loop
// (Source/Core/Shaders/Intersectors/Include/TraverseBVHStackless.glsl:175-366,
// :463-579) -- restructured as a per-lane state machine so that a 64-lane wave
// never idles on finished rays or on a few lanes' triangle tests:
//
//   * refill: when at least REFILL lanes of a wave hold no ray, one lane claims
//     that many ray indices with a single atomicAdd on a global counter and the
//     idle lanes take consecutive indices by their rank in the ballot mask
//     (mbcnt), so a wave stays full until the batch is exhausted;
//   * node phase: lanes step through nodes; a lane whose box test enters a
//     leaf parks there (its leaf must be tested before its next box test,
//     because the hit shortens TMax);
//   * leaf phase: once enough lanes are parked (or none can step) the parked
//     lanes test their leaf triangles together, both triangles of a 2-leaf
//     fetched up front.
//
// Each ray still performs the reference's operations in the reference's order
// (same node sequence, same triangle order, same TMax updates), so hit records
// are bit-identical to the one-ray-per-thread kernel and to the host oracle.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"
#include "trace_launch.h"

#pragma clang fp contract(off)

namespace crt {
namespace pdev {

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 neg(f3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
__device__ __forceinline__ float mn(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float mx(float a, float b) { return fmaxf(a, b); }

__device__ __forceinline__ f3 xform(const float* m, f3 p, float w) {
    f3 r;
    r.x = m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12] * w;
    r.y = m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13] * w;
    r.z = m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14] * w;
    return r;
}

__device__ __forceinline__ float tri_t(f3 ro, f3 rd, f3 v0, f3 v1, f3 v2) {
    f3 v1v0 = sub(v1, v0);
    f3 v2v0 = sub(v2, v0);
    f3 rov0 = sub(ro, v0);
    f3 n = cross(v1v0, v2v0);
    f3 q = cross(rov0, rd);
    float d = 1.0f / dot(rd, n);
    float u = d * dot(neg(q), v2v0);
    float v = d * dot(q, v1v0);
    float t = d * dot(neg(n), rov0);
    if (u < 0.0f || v < 0.0f || (u + v) > 1.0f) t = -1.0f;
    return t;
}

__device__ __forceinline__ float bounds_stackless(float4 bmin, float4 bmax, f3 ro, f3 inv, float t_max) {
    float t0x = (bmin.x - ro.x) * inv.x, t0y = (bmin.y - ro.y) * inv.y, t0z = (bmin.z - ro.z) * inv.z;
    float t1x = (bmax.x - ro.x) * inv.x, t1y = (bmax.y - ro.y) * inv.y, t1z = (bmax.z - ro.z) * inv.z;
    float tmin = mx(mx(mx(mn(t0x, t1x), mn(t0y, t1y)), mn(t0z, t1z)), 0.0001f);
    float tmax = mn(mn(mx(t0x, t1x), mn(mx(t0y, t1y), mx(t0z, t1z))), t_max);
    return (tmax >= tmin) ? tmin : -1.0f;
}

__device__ __forceinline__ f3 barycentrics(f3 p, f3 a, f3 b, f3 c) {
    f3 v0 = sub(b, a), v1 = sub(c, a), v2 = sub(p, a);
    float d00 = dot(v0, v0);
    float d01 = dot(v0, v1);
    float d11 = dot(v1, v1);
    float d20 = dot(v2, v0);
    float d21 = dot(v2, v1);
    float denom = d00 * d11 - d01 * d01;
    float v = (d11 * d20 - d01 * d21) / denom;
    float w = (d00 * d21 - d01 * d20) / denom;
    float u = 1.0f - v - w;
    return mk(u, v, w);
}

__device__ __forceinline__ f3 vpos(const crt_vertex* __restrict__ verts, int32_t i) {
    const float4 p = *reinterpret_cast<const float4*>(verts[i].position);
    return mk(p.x, p.y, p.z);
}

__device__ __forceinline__ unsigned lane_rank(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

__device__ __forceinline__ uint32_t pack_status(uint32_t visits, uint32_t flags) {
    return (visits > CRT_STATUS_VISITS_MASK ? CRT_STATUS_VISITS_MASK : visits) | flags;
}

// Per-lane state of one ray in the scene loop.
struct Lane {
    int32_t slot;      // ray index (a launch holds < 2^31 rays)
    f3 lro, lrd, inv;  // ray in the current entity's space
    float ray_tmax;
    int32_t ent;       // current entity
    int32_t start, count;
    int32_t ptr, it;
    float wtmax, wt;   // walk TMax and ClosestTraversal
    int32_t wtri, wmesh;
    float stmax, closest;  // scene loop TMax / ClosestT
    int32_t ctri, cmesh, cent;
    uint32_t visits, status, tests;
    int32_t leaf_packed, leaf_link;
};

template <bool ANY>
__device__ __forceinline__ void begin_walk(const SceneArgs& s, const crt_ray* __restrict__ rays, Lane& L,
                                           bool& walk_done, bool& finished, uint32_t flags) {
    // Advance to the next entity that takes part (IgnoreTransparent skips alpha < 0.99).
    int32_t e = L.ent + 1;
    while (e < s.entity_count && (flags & CRT_FLAG_IGNORE_TRANSPARENT) &&
           __int_as_float(s.entities[e].data[1]) < 0.99f)
        ++e;
    L.ent = e;
    if (e >= s.entity_count) {
        finished = true;
        return;
    }
    const crt_entity& en = s.entities[e];
    const float4* rp = reinterpret_cast<const float4*>(rays + L.slot);
    const float4 a = rp[0], b = rp[1];
    const f3 ro = mk(a.x, a.y, a.z), rd = mk(b.x, b.y, b.z);
    L.lro = xform(en.inverse, ro, 1.0f);
    L.lrd = xform(en.inverse, rd, 0.0f);
    L.inv = mk(1.0f / L.lrd.x, 1.0f / L.lrd.y, 1.0f / L.lrd.z);
    L.start = en.node_offset;
    L.count = en.node_count;
    L.ptr = L.start;
    L.it = 0;
    L.wtmax = ANY ? L.ray_tmax : L.stmax;
    L.wt = -1.0f;
    L.wtri = -1;
    L.wmesh = -1;
    walk_done = L.ptr < 0;  // `while (Pointer >= 0 ...)` never runs
}

template <bool ANY>
__device__ __forceinline__ void finish_ray(const SceneArgs& s, const crt_ray* __restrict__ rays,
                                           crt_hit* __restrict__ hits, uint32_t* __restrict__ tri_tests,
                                           const Lane& L, uint32_t flags) {
    float4 tuvw = make_float4(-1.0f, -1.0f, -1.0f, -1.0f);
    int32_t tri = -1, mesh = -1, ent = -1;
    if (ANY) {
        if (L.closest > 0.0f) {
            tuvw.x = L.closest;
            tri = L.ctri;
            mesh = L.cmesh;
            ent = L.cent;
        }
    } else {
        tri = L.ctri;
        mesh = L.cmesh;
        ent = L.cent;
        const bool ok = L.closest > 0.0f && ((flags & CRT_FLAG_TRI0_IS_MISS) ? tri > 0 : tri >= 0);
        if (ok) {
            const crt_entity& e = s.entities[ent];
            const float4* rp = reinterpret_cast<const float4*>(rays + L.slot);
            const float4 a = rp[0], b = rp[1];
            const f3 lro = xform(e.inverse, mk(a.x, a.y, a.z), 1.0f);
            const f3 lrd = xform(e.inverse, mk(b.x, b.y, b.z), 0.0f);
            const int4 tr = *reinterpret_cast<const int4*>(s.tris[tri].packed);
            const f3 va = vpos(s.verts, tr.x), vb = vpos(s.verts, tr.y), vc = vpos(s.verts, tr.z);
            const f3 p = mk(lro.x + lrd.x * L.closest, lro.y + lrd.y * L.closest, lro.z + lrd.z * L.closest);
            const f3 uvw = barycentrics(p, va, vb, vc);
            tuvw = make_float4(L.closest, uvw.x, uvw.y, uvw.z);
        }
    }
    float4* o = reinterpret_cast<float4*>(hits + L.slot);
    o[0] = tuvw;
    o[1] = make_float4(__int_as_float(tri), __int_as_float(mesh), __int_as_float(ent),
                       __uint_as_float(pack_status(L.visits, L.status)));
    if (tri_tests) tri_tests[L.slot] = L.tests;
}

// Tests triangle `idx` (already fetched) against the lane's walk.  Returns true
// when an any-hit walk has its answer.
template <bool ANY>
__device__ __forceinline__ bool test_one(Lane& L, int32_t idx, int4 tr, f3 a, f3 b, f3 c) {
    const float t = tri_t(L.lro, L.lrd, a, b, c);
    L.tests++;
    if (t > 0.0f && t < L.wtmax) {
        L.wtmax = t;
        L.wt = t;
        L.wmesh = tr.w;
        L.wtri = idx;
        if (ANY) return true;
    }
    return false;
}

template <bool ANY, int REFILL, int INNER, int LEAF_BATCH>
__global__ void __launch_bounds__(256) persistent_stackless(SceneArgs s, const crt_ray* __restrict__ rays,
                                                            crt_hit* __restrict__ hits, int64_t n, uint32_t flags,
                                                            int32_t max_iter,
                                                            unsigned long long* __restrict__ counter,
                                                            uint32_t* __restrict__ tri_tests) {
    const float4* __restrict__ nodes = reinterpret_cast<const float4*>(s.nodes);
    Lane L;
    L.slot = 0;
    bool active = false, parked = false, walk_done = false;

    while (true) {
        // --- refill idle lanes from the global ray counter ---------------------
        const unsigned long long idle = __ballot(!active);
        if (idle) {
            const int nidle = __popcll(idle);
            if (idle == ~0ull || nidle >= REFILL) {
                unsigned long long base = 0;
                if (__lane_id() == 0) base = atomicAdd(counter, (unsigned long long)nidle);
                base = (unsigned long long)__shfl((long long)base, 0);
                if (base >= (unsigned long long)n && idle == ~0ull) break;  // batch drained
                if (!active) {
                    const int64_t my = int64_t(base) + int64_t(lane_rank(idle));
                    if (my < n) {  // n < 2^31, so `my` fits the 32-bit slot
                        active = true;
                        L.slot = int32_t(my);
                        const float4 a = reinterpret_cast<const float4*>(rays + my)[0];
                        L.ray_tmax = a.w;
                        L.ent = -1;
                        L.stmax = a.w;
                        L.closest = -1.0f;
                        L.ctri = L.cmesh = L.cent = -1;
                        L.visits = L.status = L.tests = 0;
                        bool fin = false;
                        begin_walk<ANY>(s, rays, L, walk_done, fin, flags);
                        parked = false;
                        if (fin) {
                            finish_ray<ANY>(s, rays, hits, tri_tests, L, flags);
                            active = false;
                        }
                    }
                }
            }
        }
        if (__ballot(active) == 0) continue;  // nothing claimed this round; the counter decides next time

        // --- node phase ------------------------------------------------------------
        for (int k = 0; k < INNER; ++k) {
            if (active && !parked && !walk_done) {
                if (L.it >= max_iter) {
                    L.status |= CRT_STATUS_ITER_CAP;
                    walk_done = true;
                } else if (L.ptr < L.start || L.ptr > L.start + L.count || int64_t(L.ptr) >= s.total_nodes) {
                    L.status |= CRT_STATUS_BAD_INDEX;
                    walk_done = true;
                } else {
                    L.it++;
                    const float4 nmin = nodes[2 * int64_t(L.ptr)];
                    const float4 nmax = nodes[2 * int64_t(L.ptr) + 1];
                    const float tb = bounds_stackless(nmin, nmax, L.lro, L.inv, L.wtmax);
                    const int32_t link = __float_as_int(nmax.w);
                    const int32_t packed = __float_as_int(nmin.w);
                    if (tb > 0.0f && tb < L.wtmax) {
                        if (packed != -1) {
                            parked = true;
                            L.leaf_packed = packed;
                            L.leaf_link = link;
                        } else {
                            L.ptr++;
                        }
                    } else if (link < 0) {
                        walk_done = true;
                    } else {
                        L.ptr = link + L.start;
                    }
                }
            }
            const unsigned long long stepping = __ballot(active && !parked && !walk_done);
            if (stepping == 0 || __popcll(__ballot(parked)) >= LEAF_BATCH) break;
        }

        // --- leaf phase --------------------------------------------------------------
        if (__ballot(parked)) {
            if (parked) {
                const int32_t first = L.leaf_packed >> 4;
                const int32_t len = L.leaf_packed & 0xF;
                bool done = false;
                // Both triangles of a two-triangle leaf are fetched before either test
                // (a zero-length leaf, e.g. an empty object's root, reads nothing).
                if (len > 0) {
                    const int4 t0 = *reinterpret_cast<const int4*>(s.tris[first].packed);
                    int4 t1 = t0;
                    if (len > 1) t1 = *reinterpret_cast<const int4*>(s.tris[first + 1].packed);
                    const f3 a0 = vpos(s.verts, t0.x), b0 = vpos(s.verts, t0.y), c0 = vpos(s.verts, t0.z);
                    f3 a1 = a0, b1 = b0, c1 = c0;
                    if (len > 1) {
                        a1 = vpos(s.verts, t1.x);
                        b1 = vpos(s.verts, t1.y);
                        c1 = vpos(s.verts, t1.z);
                    }
                    done = test_one<ANY>(L, first, t0, a0, b0, c0);
                    if (!done && len > 1) done = test_one<ANY>(L, first + 1, t1, a1, b1, c1);
                    for (int32_t idx = first + 2; !done && idx < first + len; ++idx) {
                        const int4 tr = *reinterpret_cast<const int4*>(s.tris[idx].packed);
                        done = test_one<ANY>(L, idx, tr, vpos(s.verts, tr.x), vpos(s.verts, tr.y),
                                             vpos(s.verts, tr.z));
                    }
                }
                parked = false;
                if (done) {
                    walk_done = true;
                } else if (L.leaf_link < 0) {
                    walk_done = true;
                } else {
                    L.ptr = L.leaf_link + L.start;
                }
            }
        }

        // --- walk ends: next entity or the ray's result -------------------------------
        if (__ballot(walk_done)) {
            if (walk_done) {
                walk_done = false;
                L.visits += uint32_t(L.it);
                bool fin = false;
                if (ANY) {
                    if (L.wt > 0.0f) {
                        L.closest = L.wt;
                        L.ctri = L.wtri;
                        L.cmesh = L.wmesh;
                        L.cent = L.ent;
                        fin = true;
                    }
                } else if (L.wt > 0.0f && L.wt < L.stmax) {
                    L.stmax = L.wt;
                    L.closest = L.wt;
                    L.cmesh = L.wmesh;
                    L.ctri = L.wtri;
                    L.cent = L.ent;
                }
                if (!fin) begin_walk<ANY>(s, rays, L, walk_done, fin, flags);
                if (fin) {
                    finish_ray<ANY>(s, rays, hits, tri_tests, L, flags);
                    active = false;
                }
            }
        }
    }
}

__global__ void reset_counter(unsigned long long* c) { *c = 0ull; }

}  // namespace pdev

// Presets (REFILL, INNER, LEAF_BATCH).  REFILL = 65 never refills single lanes:
// a wave takes a fresh batch of 64 rays only when all of its lanes are idle,
// which keeps the leaf batching and drops the per-lane refill.
template <bool ANY>
static hipError_t launch_p(const SceneArgs& s, const crt_ray* rays, crt_hit* hits, int64_t n, uint32_t flags,
                           int32_t max_iter, unsigned long long* counter, uint32_t* tri_tests, int grid,
                           hipStream_t stream, const PersistentShape& shape) {
    constexpr int B = 256;
#define CRT_PS_LAUNCH(R, I, LB)                                                                               \
    hipLaunchKernelGGL((pdev::persistent_stackless<ANY, R, I, LB>), dim3(unsigned(grid)), dim3(B), 0, stream, s, \
                       rays, hits, n, flags, max_iter, counter, tri_tests)
    switch (shape.variant) {
        case 1: CRT_PS_LAUNCH(32, 16, 32); break;
        case 2: CRT_PS_LAUNCH(65, 16, 16); break;
        case 3: CRT_PS_LAUNCH(65, 32, 32); break;
        case 4: CRT_PS_LAUNCH(16, 32, 48); break;
        case 5: CRT_PS_LAUNCH(65, 8, 8); break;
        default: CRT_PS_LAUNCH(16, 16, 16); break;
    }
#undef CRT_PS_LAUNCH
    return hipGetLastError();
}

int persistent_fill_grid(int query) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    const void* fn = query == CRT_QUERY_ANY
                         ? reinterpret_cast<const void*>(&pdev::persistent_stackless<true, 16, 16, 16>)
                         : reinterpret_cast<const void*>(&pdev::persistent_stackless<false, 16, 16, 16>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess || per_cu <= 0) per_cu = 4;
    return cus * per_cu;
}

hipError_t launch_persistent(const SceneArgs& s, int query, const crt_ray* rays, crt_hit* hits, int64_t n,
                             uint32_t flags, int32_t max_iter, unsigned long long* counter, uint32_t* tri_tests,
                             const PersistentShape& shape, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(pdev::reset_counter, dim3(1), dim3(1), 0, stream, counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t need = (n + 255) / 256;
    const int grid = int(need < shape.grid ? need : shape.grid);
    return query == CRT_QUERY_ANY
               ? launch_p<true>(s, rays, hits, n, flags, max_iter, counter, tri_tests, grid, stream, shape)
               : launch_p<false>(s, rays, hits, n, flags, max_iter, counter, tri_tests, grid, stream, shape);
}

}  // namespace crt
