// Wavefront ray generation for the consumer passes of the reference:
//
//   diffuse / AO   DiffuseTrace.glsl:441-446 (RayOrigin = P + N * 0.05,
//                  CosWeightedHemisphere(N, hash2()), Include/Sampling.glsl:1-12)
//   specular       SpecularTrace.glsl:509-513 (origin offset mix(0.05, 0.1, ...),
//                  StochasticReflectionDirection :102-135, SampleGGXVNDF
//                  Include/Sampling.glsl:63-83)
//   probes         UpdateRadianceProbes.glsl:408-423 (grid origin, LambertBRDF
//                  :351-362 normalized)
//
// The reference builds each ray inside the pass that traces it, one pixel per
// invocation.  Here the rays are written to dense buffers so that any trace
// kernel can take them: the G-buffer is a buffer of primary closest hits, and
// pixels whose primary ray missed are dropped by a stable compaction (block
// counts, one exclusive scan, then each block writes its rays at its offset in
// input order).  Everything is FP32 without contraction and without library
// sine / cosine (sampling.h), so oracle/oracle_raygen.c reproduces every ray
// bit for bit.
//
// Memory per input record: the 32-B primary ray and the 32-B hit (read twice,
// by the count and the emit kernels), the hit triangle's three positions and
// the entity's inverse matrix (cached), 32 B per ray written.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "candela_rt.h"
#include "crt_internal.h"
#include "device_math.h"
#include "sampling.h"
#include "scan.h"
#include "trace_launch.h"

#pragma clang fp contract(off)

namespace crt {
namespace rg {

using namespace ::crt::dm;
using namespace ::crt::smp;

using scan::kBlock;
using scan::kScanBlock;
using scan::block_rank;

struct Params {
    uint32_t seed;
    float offset, ao_tmax, gi_tmax;
    float spec_lo, spec_hi, spec_tmax, rscale;
    const float* rough;
    int32_t nmesh;
};

__device__ __forceinline__ bool live(const crt_hit* __restrict__ hits, int64_t i, int64_t n) {
    return i < n && hits[i].t > 0.0f;
}

__global__ void __launch_bounds__(kBlock) count_kernel(const crt_hit* __restrict__ hits, int64_t n,
                                                       uint32_t* __restrict__ block_counts) {
    __shared__ uint32_t wave_counts[kBlock / 64];
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    uint32_t total;
    (void)block_rank(live(hits, i, n), wave_counts, total);
    if (threadIdx.x == 0) block_counts[blockIdx.x] = total;
}

struct Surface {
    f3 p, n, d;
    int32_t mesh;
    uint32_t id;
};

// Hit point and world-space face normal turned against the ray.
__device__ __forceinline__ Surface surface(const SceneArgs& s, const crt_ray& r, const crt_hit& h) {
    Surface o;
    o.d = mk(r.dir[0], r.dir[1], r.dir[2]);
    o.p = add(mk(r.org[0], r.org[1], r.org[2]), scale(o.d, h.t));
    f3 a, b, c;
    if (s.tri_records) {
        const crt_tri_record& t = s.tri_records[h.tri];
        a = mk(t.v0[0], t.v0[1], t.v0[2]);
        b = mk(t.v1[0], t.v1[1], t.v1[2]);
        c = mk(t.v2[0], t.v2[1], t.v2[2]);
    } else {
        const int4 tr = *reinterpret_cast<const int4*>(s.tris[h.tri].packed);
        const float4 pa = *reinterpret_cast<const float4*>(s.verts[tr.x].position);
        const float4 pb = *reinterpret_cast<const float4*>(s.verts[tr.y].position);
        const float4 pc = *reinterpret_cast<const float4*>(s.verts[tr.z].position);
        a = mk(pa.x, pa.y, pa.z);
        b = mk(pb.x, pb.y, pb.z);
        c = mk(pc.x, pc.y, pc.z);
    }
    const f3 no = cross(sub(b, a), sub(c, a));
    const float* inv = s.entities[h.entity].inverse;
    f3 nw = normalize(mk(inv[0] * no.x + inv[1] * no.y + inv[2] * no.z, inv[4] * no.x + inv[5] * no.y + inv[6] * no.z,
                         inv[8] * no.x + inv[9] * no.y + inv[10] * no.z));
    if (dot(nw, o.d) > 0.0f) nw = scale(nw, -1.0f);
    o.n = nw;
    o.mesh = h.mesh;
    o.id = r.id;
    return o;
}

__device__ __forceinline__ void store(crt_ray* __restrict__ out, int64_t k, f3 org, float tmax, f3 dir, uint32_t id) {
    float4* o = reinterpret_cast<float4*>(out + k);
    o[0] = make_float4(org.x, org.y, org.z, tmax);
    o[1] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(id));
}

__device__ __forceinline__ f3 reflect(f3 i, f3 n) { return sub(i, scale(n, 2.0f * dot(n, i))); }

// SampleGGXVNDF (Include/Sampling.glsl:63-83): GGX half vector about N.
__device__ __forceinline__ f3 ggx_half(f3 n, float roughness, float x, float y) {
    const float alpha = roughness * roughness;
    const float alpha2 = alpha * alpha;
    float sp, cp;
    sincos_turn(x, sp, cp);
    const float cos_t = sqrtf((1.0f - y) / (1.0f + (alpha2 - 1.0f) * y));
    const float sin_t = sqrtf(1.0f - cos_t * cos_t);
    const f3 up = fabsf(n.z) < 0.999f ? mk(0.0f, 0.0f, 1.0f) : mk(1.0f, 0.0f, 0.0f);
    const f3 tangent = normalize(cross(up, n));
    const f3 bitangent = cross(n, tangent);
    return normalize(add(add(scale(tangent, cp * sin_t), scale(bitangent, sp * sin_t)), scale(n, cos_t)));
}

template <int KIND>   // 0 diffuse + AO, 1 specular
__global__ void __launch_bounds__(kBlock) emit_kernel(SceneArgs s, Params P, const crt_ray* __restrict__ primary,
                                                      const crt_hit* __restrict__ hits, int64_t n,
                                                      const int64_t* __restrict__ offsets, crt_ray* __restrict__ out0,
                                                      crt_ray* __restrict__ out1) {
    __shared__ uint32_t wave_counts[kBlock / 64];
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const bool flag = live(hits, i, n);
    uint32_t total;
    const uint32_t rank = block_rank(flag, wave_counts, total);
    if (!flag) return;
    const int64_t k = offsets[blockIdx.x] + rank;
    const Surface sf = surface(s, primary[i], hits[i]);
    uint32_t h = pcg(sf.id ^ (P.seed * 0x9E3779B9u));
    if (KIND == 0) {
        const f3 org = add(sf.p, scale(sf.n, P.offset));
        const float u0 = unit(h);
        h = pcg(h);
        const float u1 = unit(h);
        h = pcg(h);
        const float u2 = unit(h);
        h = pcg(h);
        const float u3 = unit(h);
        if (out0) store(out0, k, org, P.ao_tmax, normalize(cosine_dir(sf.n, u0, u1)), sf.id);
        if (out1) store(out1, k, org, P.gi_tmax, normalize(cosine_dir(sf.n, u2, u3)), sf.id);
    } else {
        const float r = (P.rough && sf.mesh >= 0 && sf.mesh < P.nmesh) ? P.rough[sf.mesh] : 1.0f;
        const float t = fminf(fmaxf(r * 1.4f, 0.0f), 1.0f);
        const float off = P.spec_lo * (1.0f - t) + P.spec_hi * t;
        const f3 org = add(sf.p, scale(sf.n, off));
        const f3 inc = normalize(sf.d);
        const float rr = r * P.rscale;
        f3 m = sf.n;
        if (!(rr < 0.01f)) {
            for (int tries = 0; tries < 12; ++tries) {
                const float x = unit(h) * 0.8f;
                h = pcg(h);
                const float y = unit(h) * 0.7f;
                h = pcg(h);
                const f3 hv = ggx_half(sf.n, rr, x, y);
                if (dot(hv, sf.n) > 0.001f) {
                    m = hv;
                    break;
                }
            }
        }
        store(out0, k, org, P.spec_tmax, reflect(inc, m), sf.id);
    }
}

__global__ void __launch_bounds__(kBlock) probe_kernel(crt_probe_grid g, int64_t total, crt_ray* __restrict__ out) {
    const int64_t p = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (p >= total) return;
    const int64_t rx = g.res[0], ry = g.res[1];
    const int32_t x = int32_t(p % rx), y = int32_t((p / rx) % ry), z = int32_t(p / (rx * ry));
    const float cx = (float(x) / float(g.res[0])) * 2.0f - 1.0f;
    const float cy = (float(y) / float(g.res[1])) * 2.0f - 1.0f;
    const float cz = (float(z) / float(g.res[2])) * 2.0f - 1.0f;
    const f3 org = mk(g.box_origin[0] + cx * g.size[0], g.box_origin[1] + cy * g.size[1],
                      g.box_origin[2] + cz * g.size[2]);
    uint32_t h = pcg(uint32_t(p) ^ (g.seed * 0x9E3779B9u));
    const float u0 = unit(h);
    h = pcg(h);
    const float u1 = unit(h);
    float sp, cp;
    sincos_turn(u0, sp, cp);
    const float ct = 2.0f * u1 - 1.0f;
    const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    store(out, p, org, g.tmax, normalize(mk(st * cp, st * sp, ct)), uint32_t(p));
}

inline int64_t blocks_of(int64_t n) { return (n + kBlock - 1) / kBlock; }
inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

int generate(int kind, const char* who, const crt_scene_view* scene, const crt_ray* primary, const crt_hit* hits,
             int64_t n, const crt_raygen_params* params, crt_ray* out0, crt_ray* out1, int64_t* d_count,
             void* workspace, size_t workspace_bytes, void* stream) {
    const std::string w(who);
    if (!scene || !params || !d_count) return set_error(CRT_E_INVALID, w + ": null argument");
    if (n < 0 || (n > 0 && (!primary || !hits))) return set_error(CRT_E_INVALID, w + ": bad input buffers");
    if (kind == 1 && !out0) return set_error(CRT_E_INVALID, w + ": out is null");
    if (params->nmesh < 0 || (params->nmesh > 0 && !params->roughness))
        return set_error(CRT_E_INVALID, w + ": bad roughness table");
    const int64_t nb = blocks_of(n);
    if (nb > 0x7FFFFFFF) return set_error(CRT_E_INVALID, w + ": too many records");
    if (n > 0 && (!scene->entities || (!scene->tri_records && (!scene->triangles || !scene->vertices))))
        return set_error(CRT_E_INVALID, w + ": the scene needs entities and triangle positions");
    const size_t need = crt_raygen_workspace_bytes(n);
    if (!workspace || workspace_bytes < need || reinterpret_cast<uintptr_t>(workspace) % 8 != 0)
        return set_error(CRT_E_INVALID, w + ": workspace must be 8-byte aligned and hold crt_raygen_workspace_bytes(n) = " +
                                            std::to_string(need) + " bytes");
    auto* counts = static_cast<uint32_t*>(workspace);
    auto* offsets = reinterpret_cast<int64_t*>(static_cast<char*>(workspace) + align256(size_t(nb) * 4));
    const hipStream_t st = static_cast<hipStream_t>(stream);
    Params P{params->seed,          params->offset,         params->ao_tmax,   params->gi_tmax,
             params->spec_offset[0], params->spec_offset[1], params->spec_tmax, params->roughness_scale,
             params->roughness,     params->nmesh};
    const SceneArgs a = scene_args(*scene);
    if (nb > 0) hipLaunchKernelGGL(count_kernel, dim3(unsigned(nb)), dim3(kBlock), 0, st, hits, n, counts);
    hipLaunchKernelGGL(scan::exclusive_scan, dim3(1), dim3(kScanBlock), 0, st, counts, nb, offsets, d_count);
    if (nb > 0) {
        if (kind == 0)
            hipLaunchKernelGGL(emit_kernel<0>, dim3(unsigned(nb)), dim3(kBlock), 0, st, a, P, primary, hits, n,
                               offsets, out0, out1);
        else
            hipLaunchKernelGGL(emit_kernel<1>, dim3(unsigned(nb)), dim3(kBlock), 0, st, a, P, primary, hits, n,
                               offsets, out0, out1);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(CRT_E_HIP, w + ": kernel launch: " + hipGetErrorString(e));
    return CRT_OK;
}

}  // namespace rg
}  // namespace crt

extern "C" {

size_t crt_raygen_workspace_bytes(int64_t n) {
    const int64_t nb = n > 0 ? crt::rg::blocks_of(n) : 0;
    return crt::rg::align256(size_t(nb) * 4) + size_t(nb) * 8 + 256;
}

int crt_gen_diffuse_rays(const crt_scene_view* scene, const crt_ray* primary, const crt_hit* hits, int64_t n,
                         const crt_raygen_params* params, crt_ray* ao_out, crt_ray* gi_out, int64_t* d_count,
                         void* workspace, size_t workspace_bytes, void* stream) {
    return crt::rg::generate(0, "crt_gen_diffuse_rays", scene, primary, hits, n, params, ao_out, gi_out, d_count,
                             workspace, workspace_bytes, stream);
}

int crt_gen_specular_rays(const crt_scene_view* scene, const crt_ray* primary, const crt_hit* hits, int64_t n,
                          const crt_raygen_params* params, crt_ray* out, int64_t* d_count, void* workspace,
                          size_t workspace_bytes, void* stream) {
    return crt::rg::generate(1, "crt_gen_specular_rays", scene, primary, hits, n, params, out, nullptr, d_count,
                             workspace, workspace_bytes, stream);
}

int crt_gen_probe_rays(const crt_probe_grid* grid, crt_ray* out, void* stream) {
    if (!grid || !out) return crt::set_error(CRT_E_INVALID, "crt_gen_probe_rays: null argument");
    if (grid->res[0] <= 0 || grid->res[1] <= 0 || grid->res[2] <= 0)
        return crt::set_error(CRT_E_INVALID, "crt_gen_probe_rays: grid resolution must be positive");
    const int64_t total = int64_t(grid->res[0]) * grid->res[1] * grid->res[2];
    if (total > 0xFFFFFFFFll) return crt::set_error(CRT_E_INVALID, "crt_gen_probe_rays: more than 2^32 probes");
    const int64_t nb = crt::rg::blocks_of(total);
    hipLaunchKernelGGL(crt::rg::probe_kernel, dim3(unsigned(nb)), dim3(crt::rg::kBlock), 0,
                       static_cast<hipStream_t>(stream), *grid, total, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return crt::set_error(CRT_E_HIP, std::string("crt_gen_probe_rays: kernel launch: ") + hipGetErrorString(e));
    return CRT_OK;
}

}  // extern "C"
