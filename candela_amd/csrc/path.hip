// Multi-bounce diffuse path tracing with wavefront path regeneration in one
// persistent kernel (the large-scene workload: a displaced terrain plus
// icospheres, 3840 x 2160 x 4 samples, 2 diffuse bounces, Russian roulette
// after the first bounce).
//
// Every lane owns one path.  A path's segments are walked by the wide walk of
// wide_walk.h (Core, same node step, same pending-leaf parking); when a
// segment's entity loop ends, the lane shades in place -- a miss adds
// throughput x sky and ends the path, a hit multiplies the throughput by the
// mesh albedo and, unless the path is done, spawns the next segment -- and
// keeps stepping.  A lane whose path ended claims the next path index from
// the sharded queue (queue.h) together with the other idle lanes of its wave
// (one atomic per claim, ranks by mbcnt) and generates that path's camera ray
// on the spot, so no ray or hit buffer exists between segments and no wave
// waits for its longest path.
//
// Per path p (all FP32, no contraction; the host twin in oracle/oracle_path.c
// follows the same sequence bit for bit):
//   rng      state = pcg(p ^ (seed * 0x9E3779B9)); each draw state = pcg(state),
//            uniform = (state >> 8) * 2^-24
//   camera   pixel = p / spp, x = pixel % width, y = pixel / width; jitter
//            (draw, draw); IntersectPrimary's GetRayDirectionAt
//            (TraverseBVHStackless.glsl:143-148) at uv = ((x + jx) / width,
//            (y + jy) / height), direction normalised
//   segment  closest hit over the scene (entity loop, tmax as given)
//   miss     radiance += throughput * sky; end
//   hit      throughput *= albedo[mesh] (0 for a mesh beyond the table);
//            end after `max_bounces` bounces; from the second bounce on,
//            continue with probability rr (draw < rr) and throughput /= rr;
//            next direction cosine-weighted about the face normal turned
//            against the ray (basis of CosWeightedHemisphere,
//            Include/Sampling.glsl:1-12; sin and cos of the azimuth from a
//            quarter-angle Taylor polynomial so host and device agree),
//            origin = hit + normal * 1e-4
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "candela_rt.h"
#include "crt_internal.h"
#include "trace_launch.h"
#include "sampling.h"
#include "spec_walk.h"
#include "wide_walk.h"

#pragma clang fp contract(off)

namespace crt {
namespace path {

using namespace ::crt::dm;
using wide::Core;
using wide::kEnd;
using wide::kLeaves;
using wide::Stack;
using namespace ::crt::smp;

__device__ __forceinline__ float draw(uint32_t& st) {
    st = pcg(st);
    return unit(st);
}

struct Params {
    Camera cam;
    int32_t width, height, spp;
    uint32_t seed;
    int32_t max_bounces;
    float rr;
    float sky[3];
    const float* albedo;   // nmesh * 3
    int32_t nmesh;
};

template <int W, int LDS_DEPTH, int INNER, int LEAF_BATCH>
struct PathWave {
    Core<false, W, LDS_DEPTH> c;
    int64_t path;
    f3 org, dir;
    f3 thr, rad;
    uint32_t rng;
    int32_t bounce, segs;
    uint32_t visits, tests;
    int32_t ent, hit_ent;
    int32_t leaf, off;
    bool active, leafing, walk_done;

    __device__ __forceinline__ bool begin_walk(const SceneArgs& s) {
        const int32_t e = ent + 1;
        ent = e;
        if (e >= s.entity_count) return false;
        const crt_entity& en = s.entities[e];
        c.begin(xform(en.inverse, org, 1.0f), xform(en.inverse, dir, 0.0f), c.tmax, en.node_offset, en.node_count,
                s.total_nodes);
        walk_done = false;
        return true;
    }

    // A new segment from (org, dir); false when the scene has no entity.
    __device__ __forceinline__ bool start_segment(const SceneArgs& s, float tmax) {
        ent = -1;
        hit_ent = -1;
        c.tmax = tmax;
        c.tri = c.mesh = -1;
        c.tests = 0;
        segs++;
        return begin_walk(s);
    }

    // The segment's entity loop is over: shade.  True when the path continues
    // (org, dir hold the next segment's ray).
    __device__ __forceinline__ bool shade(const SceneArgs& s, const Params& P) {
        tests += c.tests;
        if (c.tri < 0) {
            rad = add(rad, mk(thr.x * P.sky[0], thr.y * P.sky[1], thr.z * P.sky[2]));
            return false;
        }
        const float t = c.tmax;
        const int32_t mesh = c.mesh;
        f3 alb = mk(0.0f, 0.0f, 0.0f);
        if (mesh >= 0 && mesh < P.nmesh) alb = mk(P.albedo[3 * mesh], P.albedo[3 * mesh + 1], P.albedo[3 * mesh + 2]);
        thr = mk(thr.x * alb.x, thr.y * alb.y, thr.z * alb.z);
        if (bounce >= P.max_bounces) return false;
        if (bounce >= 1) {
            if (!(draw(rng) < P.rr)) return false;
            thr = mk(thr.x / P.rr, thr.y / P.rr, thr.z / P.rr);
        }
        // face normal of the hit triangle in world space (inverse-transpose of the entity's model)
        const float4* r = reinterpret_cast<const float4*>(s.tri_records) + 3 * int64_t(c.tri);
        const float4 va = r[0], vb = r[1], vc = r[2];
        const f3 a = mk(va.x, va.y, va.z);
        const f3 no = cross(sub(mk(vb.x, vb.y, vb.z), a), sub(mk(vc.x, vc.y, vc.z), a));
        const float* inv = s.entities[hit_ent].inverse;
        f3 n = normalize(mk(inv[0] * no.x + inv[1] * no.y + inv[2] * no.z, inv[4] * no.x + inv[5] * no.y + inv[6] * no.z,
                            inv[8] * no.x + inv[9] * no.y + inv[10] * no.z));
        if (dot(n, dir) > 0.0f) n = neg(n);
        const f3 p = add(org, scale(dir, t));
        const float u1 = draw(rng), u2 = draw(rng);
        org = add(p, scale(n, 1e-4f));
        dir = cosine_dir(n, u1, u2);
        bounce++;
        return true;
    }

    __device__ __forceinline__ void finish(crt_path_sample* __restrict__ out) {
        float4* o = reinterpret_cast<float4*>(out + path);
        o[0] = make_float4(rad.x, rad.y, rad.z, __int_as_float(segs));
        o[1] = make_float4(__uint_as_float(visits), __uint_as_float(tests), __uint_as_float(c.status), 0.0f);
        active = false;
    }

    // Segment over: shade, then the next segment or the path's record.
    __device__ __forceinline__ void segment_done(const SceneArgs& s, const Params& P, float tmax,
                                                 crt_path_sample* __restrict__ out) {
        while (shade(s, P)) {
            if (start_segment(s, tmax)) return;
            c.tri = -1;   // no entity: the new segment misses
        }
        finish(out);
    }

    __device__ __forceinline__ void start_path(const SceneArgs& s, const Params& P, int64_t p, float tmax,
                                               crt_path_sample* __restrict__ out) {
        active = true;
        leafing = false;
        walk_done = false;
        path = p;
        rng = pcg(uint32_t(p) ^ (P.seed * 0x9E3779B9u));
        const int64_t pixel = p / P.spp;
        const int32_t x = int32_t(pixel % P.width), y = int32_t(pixel / P.width);
        const float jx = draw(rng), jy = draw(rng);
        const float u = (float(x) + jx) / float(P.width), v = (float(y) + jy) / float(P.height);
        const float cx = u * 2.0f - 1.0f, cy = v * 2.0f - 1.0f;
        const float* ip = P.cam.inv_proj;
        const float ex = ip[0] * cx + ip[4] * cy + ip[8] * -1.0f + ip[12] * 1.0f;
        const float ey = ip[1] * cx + ip[5] * cy + ip[9] * -1.0f + ip[13] * 1.0f;
        const f3 d = xform(P.cam.inv_view, mk(ex, ey, -1.0f), 0.0f);
        dir = scale(d, 1.0f / sqrtf(dot(d, d)));
        org = mk(P.cam.inv_view[12], P.cam.inv_view[13], P.cam.inv_view[14]);
        thr = mk(1.0f, 1.0f, 1.0f);
        rad = mk(0.0f, 0.0f, 0.0f);
        bounce = 0;
        segs = 0;
        visits = 0;
        tests = 0;
        c.status = 0;
        if (!start_segment(s, tmax)) {
            c.tri = -1;
            segment_done(s, P, tmax, out);
        }
    }

    __device__ __forceinline__ void round(const SceneArgs& s, const Params& P, float tmax, int32_t max_iter,
                                          Stack<LDS_DEPTH>& st, crt_path_sample* __restrict__ out) {
        const float4* __restrict__ rec = reinterpret_cast<const float4*>(s.tri_records);
        for (int k = 0; k < INNER; ++k) {
            if (active && !walk_done && !leafing) {
                const int32_t r = c.step(s, max_iter, st, false);
                if (r == kEnd) {
                    walk_done = true;
                } else if (r == kLeaves) {
                    leafing = true;
                    leaf = st.leaf(0);
                    c.pi = 1;
                    off = 0;
                }
            }
            const unsigned long long stepping = __ballot(active && !walk_done && !leafing);
            if (stepping == 0 || __popcll(__ballot(leafing)) >= LEAF_BATCH) break;
        }
        while (__ballot(leafing)) {
            if (leafing) {
                const int32_t first = (leaf >> 4) + off, left = (leaf & 0xF) - off;
                const float4* p = rec + 3 * int64_t(first);
                // the second record, or the first again when one is left
                const float4* p1 = p + (left > 1 ? 3 : 0);
                const float4 a0 = p[0], b0 = p[1], c0 = p[2];
                const float4 a1 = p1[0], b1 = p1[1], c1 = p1[2];
                const float before = c.tmax;
                c.test(first, a0, b0, c0);
                if (left > 1) c.test(first + 1, a1, b1, c1);
                if (c.tmax < before) hit_ent = ent;
                off += left > 1 ? 2 : 1;
                if (off >= (leaf & 0xF)) {
                    if (c.pi < c.np) {
                        leaf = st.leaf(c.pi);
                        ++c.pi;
                        off = 0;
                    } else {
                        leafing = false;
                    }
                }
            }
        }
        if (__ballot(walk_done)) {
            if (walk_done) {
                walk_done = false;
                visits += uint32_t(c.it);
                if (!begin_walk(s)) segment_done(s, P, tmax, out);
            }
        }
    }

    __device__ __forceinline__ void run(const SceneArgs& s, const Params& P, int64_t npaths, float tmax,
                                        int32_t max_iter, Stack<LDS_DEPTH> st, unsigned long long* __restrict__ counters,
                                        int32_t refill, crt_path_sample* __restrict__ out) {
        active = false;
        leafing = false;
        walk_done = false;
        ShardQueue q;
        q.init(counters, npaths);
        while (true) {
            if (q.more) {
                const unsigned long long idle = __ballot(!active);
                const int32_t cnt = __popcll(idle);
                if (cnt >= refill) {
                    int32_t got = 0;
                    const int64_t first = q.claim(cnt, got);
                    if (!active) {
                        const int32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(idle >> 32),
                                                                       __builtin_amdgcn_mbcnt_lo(uint32_t(idle), 0));
                        if (rank < got) start_path(s, P, q.at(first + rank), tmax, out);
                    }
                }
            }
            if (!__ballot(active)) {
                if (!q.more) break;
                continue;
            }
            round(s, P, tmax, max_iter, st, out);
        }
    }
};

template <int W, int LDS_DEPTH, int INNER, int LEAF_BATCH, int REFILL>
__device__ __forceinline__ void path_body(const SceneArgs& s, const Params& P, int64_t npaths, float tmax,
                                          int32_t max_iter, unsigned long long* __restrict__ counters,
                                          crt_path_sample* __restrict__ out, int32_t* lds) {
    int32_t spill_node[wide::kSpill];
    float spill_t[wide::kSpill];
    Stack<LDS_DEPTH> st;
    st.lds = lds + threadIdx.x;
    st.stride = 256;
    st.spill_node = spill_node;
    st.spill_t = spill_t;
    PathWave<W, LDS_DEPTH, INNER, LEAF_BATCH> w;
    w.run(s, P, npaths, tmax, max_iter, st, counters, REFILL, out);
}

template <int W, int LDS_DEPTH, int INNER, int LEAF_BATCH, int REFILL>
__global__ void __launch_bounds__(256) path_kernel(SceneArgs s, Params P, int64_t npaths, float tmax,
                                                   int32_t max_iter, unsigned long long* __restrict__ counters,
                                                   crt_path_sample* __restrict__ out) {
    __shared__ int32_t lds[wide::lds_ints<W, LDS_DEPTH>() * 256];
    path_body<W, LDS_DEPTH, INNER, LEAF_BATCH, REFILL>(s, P, npaths, tmax, max_iter, counters, out, lds);
}

// The same kernel compiled for at least OCC waves per SIMD (register budget 512 / OCC).
template <int W, int LDS_DEPTH, int INNER, int LEAF_BATCH, int REFILL, int OCC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC, 8)))
path_kernel_occ(SceneArgs s, Params P, int64_t npaths, float tmax, int32_t max_iter,
                unsigned long long* __restrict__ counters, crt_path_sample* __restrict__ out) {
    __shared__ int32_t lds[wide::lds_ints<W, LDS_DEPTH>() * 256];
    path_body<W, LDS_DEPTH, INNER, LEAF_BATCH, REFILL>(s, P, npaths, tmax, max_iter, counters, out, lds);
}

// The same paths over the speculative walk (spec_walk.h): a segment's leaves go
// to the lane's queue and the wave tests every queued leaf once per round, so
// no lane waits for its leaf tests.  The walk state is the SpecWave's (its ray
// fields hold the segment's ray in the current entity's space, `active` marks
// a live path); the path state and shading are PathWave's.  Each segment ends
// with the nearest hit the sequential walk finds, so radiance and segment
// counts equal the host twin's; visit and test counts follow the schedule.
// Segment ends of the speculative path walk wait for this many lanes (or for
// a round in which no lane can step): measured on the large scene, 24 takes a
// frame from 12.64 to 12.03 ms; 16 gives 12.14, 32 12.13 and 48 12.44.
constexpr int32_t kPathEndBatch = 24;

template <int W, int LDS_DEPTH, int INNER, int Q>
struct PathSpecWave {
    spec::SpecWave<false, W, LDS_DEPTH, INNER, Q> w;
    int64_t path;
    f3 org, dir;
    f3 thr, rad;
    uint32_t rng;
    int32_t bounce, segs;
    uint32_t visits, tests;

    __device__ __forceinline__ bool begin_walk(const SceneArgs& s) {
        const int32_t e = w.ent + 1;
        w.ent = e;
        if (e >= s.entity_count) return false;
        const crt_entity& en = s.entities[e];
        w.begin_local(xform(en.inverse, org, 1.0f), xform(en.inverse, dir, 0.0f), en.node_offset, en.node_count,
                      s.total_nodes);
        return true;
    }

    __device__ __forceinline__ bool start_segment(const SceneArgs& s, float tmax) {
        w.ent = -1;
        w.hit_ent = -1;
        w.tmax = tmax;
        w.tri = w.mesh = -1;
        w.tests = 0;
        segs++;
        return begin_walk(s);
    }

    // PathWave::shade over the walk's hit.
    __device__ __forceinline__ bool shade(const SceneArgs& s, const Params& P) {
        tests += w.tests;
        if (w.tri < 0) {
            rad = add(rad, mk(thr.x * P.sky[0], thr.y * P.sky[1], thr.z * P.sky[2]));
            return false;
        }
        const float t = w.tmax;
        const int32_t mesh = w.mesh;
        f3 alb = mk(0.0f, 0.0f, 0.0f);
        if (mesh >= 0 && mesh < P.nmesh) alb = mk(P.albedo[3 * mesh], P.albedo[3 * mesh + 1], P.albedo[3 * mesh + 2]);
        thr = mk(thr.x * alb.x, thr.y * alb.y, thr.z * alb.z);
        if (bounce >= P.max_bounces) return false;
        if (bounce >= 1) {
            if (!(draw(rng) < P.rr)) return false;
            thr = mk(thr.x / P.rr, thr.y / P.rr, thr.z / P.rr);
        }
        const float4* r = reinterpret_cast<const float4*>(s.tri_records) + 3 * int64_t(w.tri);
        const float4 va = r[0], vb = r[1], vc = r[2];
        const f3 a = mk(va.x, va.y, va.z);
        const f3 no = cross(sub(mk(vb.x, vb.y, vb.z), a), sub(mk(vc.x, vc.y, vc.z), a));
        const float* inv = s.entities[w.hit_ent].inverse;
        f3 n = normalize(mk(inv[0] * no.x + inv[1] * no.y + inv[2] * no.z, inv[4] * no.x + inv[5] * no.y + inv[6] * no.z,
                            inv[8] * no.x + inv[9] * no.y + inv[10] * no.z));
        if (dot(n, dir) > 0.0f) n = neg(n);
        const f3 p = add(org, scale(dir, t));
        const float u1 = draw(rng), u2 = draw(rng);
        org = add(p, scale(n, 1e-4f));
        dir = cosine_dir(n, u1, u2);
        bounce++;
        return true;
    }

    __device__ __forceinline__ void finish(crt_path_sample* __restrict__ out) {
        float4* o = reinterpret_cast<float4*>(out + path);
        o[0] = make_float4(rad.x, rad.y, rad.z, __int_as_float(segs));
        o[1] = make_float4(__uint_as_float(visits), __uint_as_float(tests), __uint_as_float(w.status), 0.0f);
        w.active = false;
    }

    __device__ __forceinline__ void segment_done(const SceneArgs& s, const Params& P, float tmax,
                                                 crt_path_sample* __restrict__ out) {
        while (shade(s, P)) {
            if (start_segment(s, tmax)) return;
            w.tri = -1;   // no entity: the new segment misses
        }
        finish(out);
    }

    // PathWave::start_path (same camera ray and random stream).
    __device__ __forceinline__ void start_path(const SceneArgs& s, const Params& P, int64_t p, float tmax,
                                               crt_path_sample* __restrict__ out) {
        w.active = true;
        w.walk_done = false;
        w.nq = 0;
        path = p;
        rng = pcg(uint32_t(p) ^ (P.seed * 0x9E3779B9u));
        const int64_t pixel = p / P.spp;
        const int32_t x = int32_t(pixel % P.width), y = int32_t(pixel / P.width);
        const float jx = draw(rng), jy = draw(rng);
        const float u = (float(x) + jx) / float(P.width), v = (float(y) + jy) / float(P.height);
        const float cx = u * 2.0f - 1.0f, cy = v * 2.0f - 1.0f;
        const float* ip = P.cam.inv_proj;
        const float ex = ip[0] * cx + ip[4] * cy + ip[8] * -1.0f + ip[12] * 1.0f;
        const float ey = ip[1] * cx + ip[5] * cy + ip[9] * -1.0f + ip[13] * 1.0f;
        const f3 d = xform(P.cam.inv_view, mk(ex, ey, -1.0f), 0.0f);
        dir = scale(d, 1.0f / sqrtf(dot(d, d)));
        org = mk(P.cam.inv_view[12], P.cam.inv_view[13], P.cam.inv_view[14]);
        thr = mk(1.0f, 1.0f, 1.0f);
        rad = mk(0.0f, 0.0f, 0.0f);
        bounce = 0;
        segs = 0;
        visits = 0;
        tests = 0;
        w.status = 0;
        if (!start_segment(s, tmax)) {
            w.tri = -1;
            segment_done(s, P, tmax, out);
        }
    }

    __device__ __forceinline__ void round(const SceneArgs& s, const Params& P, float tmax, int32_t max_iter,
                                          Stack<LDS_DEPTH>& st, spec::WaveLds<Q>& L,
                                          crt_path_sample* __restrict__ out) {
        for (int k = 0; k < INNER; ++k) {
            if (w.can_step()) {
                if (!w.step(s, max_iter, st)) w.walk_done = true;
            }
            if (!__ballot(w.can_step())) break;
        }
        if (__ballot(w.nq > 0)) w.leaf_phase(s, st, L);
        // every queue is empty here: a segment's walk of this entity is over;
        // the segment ends (shading, the next segment's ray) run once
        // kPathEndBatch lanes wait or no lane can step
        const bool ends = w.active && w.walk_done;
        const unsigned long long em = __ballot(ends);
        if (em && (__popcll(em) >= kPathEndBatch || !__ballot(w.can_step()))) {
            if (ends) {
                visits += uint32_t(w.it);
                if (!begin_walk(s)) segment_done(s, P, tmax, out);
            }
        }
    }

    __device__ __forceinline__ void run(const SceneArgs& s, const Params& P, int64_t npaths, float tmax,
                                        int32_t max_iter, Stack<LDS_DEPTH> st, spec::WaveLds<Q>& L,
                                        unsigned long long* __restrict__ counters, int32_t refill,
                                        crt_path_sample* __restrict__ out) {
        w.active = false;
        w.walk_done = false;
        w.nq = 0;
        ShardQueue q;
        q.init(counters, npaths);
        while (true) {
            if (q.more) {
                const unsigned long long idle = __ballot(!w.active);
                const int32_t cnt = __popcll(idle);
                if (cnt >= refill) {
                    int32_t got = 0;
                    const int64_t first = q.claim(cnt, got);
                    if (!w.active) {
                        const int32_t rank = spec::lanes_below(idle);
                        if (rank < got) start_path(s, P, q.at(first + rank), tmax, out);
                    }
                }
            }
            if (!__ballot(w.active)) {
                if (!q.more) break;
                continue;
            }
            round(s, P, tmax, max_iter, st, L, out);
        }
    }
};

template <int W, int Q>
constexpr int spec_queue_slots() {
    return W == 4 ? Q : (Q + 4 < 16 ? Q + 4 : 16);
}

template <int W, int D, int Q>
struct SpecPathLds {
    int32_t cols[spec::lds_ints<D, spec_queue_slots<W, Q>()>() * 256];
    spec::WaveLds<spec_queue_slots<W, Q>()> waves[4];
};

template <int W, int LDS_DEPTH, int INNER, int Q, int REFILL, int OCC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1, 8)))
spec_path_kernel(SceneArgs s, Params P, int64_t npaths, float tmax, int32_t max_iter,
                 unsigned long long* __restrict__ counters, crt_path_sample* __restrict__ out) {
    __shared__ SpecPathLds<W, LDS_DEPTH, Q> lds;
    int32_t spill_node[wide::kSpill];
    float spill_t[wide::kSpill];
    Stack<LDS_DEPTH> st;
    st.lds = lds.cols + threadIdx.x;
    st.stride = 256;
    st.spill_node = spill_node;
    st.spill_t = spill_t;
    PathSpecWave<W, LDS_DEPTH, INNER, spec_queue_slots<W, Q>()> pw;
    pw.run(s, P, npaths, tmax, max_iter, st, lds.waves[threadIdx.x >> 6], counters, REFILL, out);
}

template <int W, int D, int I, int LB, int R, int OCC>
const void* path_kernel_ptr() {
    if constexpr (OCC == 0) {
        return reinterpret_cast<const void*>(&path_kernel<W, D, I, LB, R>);
    } else {
        return reinterpret_cast<const void*>(&path_kernel_occ<W, D, I, LB, R, OCC>);
    }
}

// Presets (variant, LDS stack entries, INNER, LEAF_BATCH, REFILL, OCC waves per
// SIMD the register budget must allow; 0 = the compiler's choice).
#define CRT_PATH_PRESETS(X) \
    X(0, 8, 4, 32, 16, 0)   \
    X(1, 8, 4, 16, 8, 0)    \
    X(2, 8, 8, 32, 32, 0)   \
    X(3, 8, 4, 48, 24, 0)   \
    X(4, 8, 4, 32, 16, 4)   \
    X(5, 8, 4, 48, 24, 4)

// Presets of the speculative path walk (variant, LDS stack entries, INNER,
// Q queue slots over the BVH4, REFILL, OCC).
#define CRT_SPEC_PATH_PRESETS(X) \
    X(6, 8, 8, 8, 16, 0)         \
    X(7, 8, 8, 8, 16, 5)         \
    X(8, 8, 4, 8, 16, 5)         \
    X(9, 8, 8, 8, 8, 5)

template <int W>
const void* path_fn(int variant) {
#define CRT_PATH_FN(V, D, I, LB, R, OCC) \
    case V: return path_kernel_ptr<W, D, I, LB, R, OCC>();
#define CRT_SPEC_PATH_FN(V, D, I, Q, R, OCC) \
    case V: return reinterpret_cast<const void*>(&spec_path_kernel<W, D, I, Q, R, OCC>);
    switch (variant) {
        CRT_PATH_PRESETS(CRT_PATH_FN)
        CRT_SPEC_PATH_PRESETS(CRT_SPEC_PATH_FN)
        default: return nullptr;
    }
#undef CRT_SPEC_PATH_FN
#undef CRT_PATH_FN
}

}  // namespace path
}  // namespace crt

extern "C" int crt_trace_paths(const crt_scene_view* scene, const crt_path_params* params, crt_path_sample* out,
                               int64_t npaths, const crt_trace_config* config, void* stream) {
    if (!scene || !params) return crt::set_error(CRT_E_INVALID, "crt_trace_paths: null scene or params");
    if (scene->format != CRT_FORMAT_WIDE4 && scene->format != CRT_FORMAT_WIDE8)
        return crt::set_error(CRT_E_INVALID, "crt_trace_paths: the path kernel walks the wide layouts");
    if (scene->entity_count > 0 && (!scene->nodes || !scene->tri_records || !scene->entities || scene->total_nodes <= 0))
        return crt::set_error(CRT_E_INVALID, "crt_trace_paths: null scene buffer");
    const crt_path_params& p = *params;
    if (p.width <= 0 || p.height <= 0 || p.spp <= 0 || p.max_bounces < 0 || !(p.rr > 0.0f && p.rr <= 1.0f))
        return crt::set_error(CRT_E_INVALID, "crt_trace_paths: bad image size, samples, bounces or roulette");
    if (p.nmesh < 0 || (p.nmesh > 0 && !p.albedo)) return crt::set_error(CRT_E_INVALID, "crt_trace_paths: bad albedo table");
    if (npaths < 0 || npaths > int64_t(p.width) * p.height * p.spp)
        return crt::set_error(CRT_E_INVALID, "crt_trace_paths: npaths exceeds width * height * spp");
    if (npaths > 0 && !out) return crt::set_error(CRT_E_INVALID, "crt_trace_paths: out is null");
    if (npaths >= (int64_t(1) << 32)) return crt::set_error(CRT_E_INVALID, "crt_trace_paths: >= 2^32 paths");
    const int variant = config ? config->variant : 0;
    if (variant < 0 || variant > 9) return crt::set_error(CRT_E_INVALID, "crt_trace_paths: variant must be 0..9");
    if (npaths == 0) return CRT_OK;
    crt::path::Params P;
    std::memcpy(P.cam.inv_view, p.inv_view, sizeof(P.cam.inv_view));
    std::memcpy(P.cam.inv_proj, p.inv_proj, sizeof(P.cam.inv_proj));
    P.cam.tmax = p.tmax;
    P.width = p.width;
    P.height = p.height;
    P.spp = p.spp;
    P.seed = p.seed;
    P.max_bounces = p.max_bounces;
    P.rr = p.rr;
    P.sky[0] = p.sky[0];
    P.sky[1] = p.sky[1];
    P.sky[2] = p.sky[2];
    P.albedo = p.albedo;
    P.nmesh = p.nmesh;
    crt::QueueLease lease;
    int rc = crt::queue_acquire(config, stream, "crt_trace_paths", lease);
    if (rc) return rc;
    unsigned long long* counters = lease.ptr;
    const void* fn = scene->format == CRT_FORMAT_WIDE8 ? crt::path::path_fn<8>(variant) : crt::path::path_fn<4>(variant);
    int grid = config && config->grid > 0 ? config->grid : 0;
    if (grid == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess || per_cu <= 0) per_cu = 4;
        grid = cus * per_cu;
    }
    const int64_t need = (npaths + 255) / 256;
    if (need < grid) grid = int(need);
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(crt::reset_queue, dim3(1), dim3(64), 0, st, counters);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        (void)crt::queue_release(lease, stream);
        return crt::set_error(CRT_E_HIP, std::string("crt_trace_paths: reset: ") + hipGetErrorString(e));
    }
    crt::SceneArgs a = crt::scene_args(*scene);
    int64_t n = npaths;
    float tmax = p.tmax;
    int32_t max_iter = crt::kMaxIterations;
    crt_path_sample* o = out;
    void* args[] = {&a, &P, &n, &tmax, &max_iter, &counters, &o};
    e = hipLaunchKernel(fn, dim3(unsigned(grid)), dim3(256), args, 0, st);
    rc = crt::queue_release(lease, stream);
    if (e != hipSuccess) return crt::set_error(CRT_E_HIP, std::string("crt_trace_paths: launch: ") + hipGetErrorString(e));
    return rc;
}
