// Differentiable shading on top of the traversal: Cook-Torrance direct light
// from point lights plus one diffuse bounce, and its hand-written adjoint.
//
// Forward, per pixel with a primary hit p (triangle t0):
//   C = sum_l CT(p, l; albedo[t0], roughness[t0]) * vis_p(l)
//     + albedo[t0] * sum_l Lambert(q, l; albedo[t1]) * vis_q(l)
// where
//   CT       CookTorranceBRDF (Source/Core/Shaders/Include/CookTorranceBRDF.glsl:37-79)
//            with metalness 0: Lo = normalize(eye - p), Li = -normalize(p - L),
//            F = 0.04 + 0.96 (1 - max(0, Lh.Lo))^5, D = NDF, G = Smith-Schlick
//            with k = (r + 1)^2 / 8, kd = 1 - F,
//            (kd * albedo + F D G / max(1e-5, 4 NL NV)) * radiance * NL;
//   Lambert  GetDirect (Source/Core/Shaders/DiffuseTrace.glsl:368-372):
//            albedo * radiance * clamp(dot(N, Li), 0, 1);
//   radiance = I_l / |L - x|^2 (white point lights);
//   q        the closest hit of one cosine-weighted bounce ray from p
//            (CosWeightedHemisphere, Source/Core/Shaders/Include/Sampling.glsl:1-12,
//            per-pixel PCG uniforms), triangle t1;
//   vis      any-hit shadow rays from x + N * 1e-4 toward L with tmax = |L - x'|.
//
// Backward, loss = sum (C - T)^2: geometry (p, q, normals) comes from the
// forward records; visibility is re-traced, not stored.  Gradients of
// albedo[t0], roughness[t0], albedo[t1] and every intensity are accumulated
// with wave-aggregated atomics: lanes that share a triangle add their
// contributions in-wave first, so the two ground-plane triangles that cover
// half the image cost one atomic per wave, not one per pixel.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"
#include "crt_internal.h"
#include "device_math.h"
#include "thread_walk.h"
#include "trace_launch.h"

namespace crt {
namespace shade {

using namespace ::crt::dm;

constexpr int kMaxLights = 16;
constexpr float kPi = 3.14159265359f;      // CookTorranceBRDF.glsl:7
constexpr float kOffset = 1e-4f;           // surface offset of secondary rays
constexpr float kTwoPi = 2.0f * 3.14159265359f;  // Sampling.glsl:3

struct Params {
    const float* albedo;
    const float* rough;
    const float* lpos;
    const float* inten;
    int32_t ntri;
    int32_t nl;
};

__device__ __forceinline__ f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 scale(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 normalize(f3 a) { return scale(a, 1.0f / sqrtf(dot(a, a))); }

__device__ __forceinline__ uint32_t pcg(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// Face normal of triangle `tri`, turned to face against `dir`.
__device__ __forceinline__ f3 face_normal(const SceneArgs& s, int32_t tri, f3 dir) {
    const int4 tr = *reinterpret_cast<const int4*>(s.tris[tri].packed);
    const f3 a = vertex_pos(s.verts, tr.x), b = vertex_pos(s.verts, tr.y), c = vertex_pos(s.verts, tr.z);
    f3 n = normalize(cross(sub(b, a), sub(c, a)));
    if (dot(n, dir) > 0.0f) n = neg(n);
    return n;
}

template <int FORMAT>
__device__ __forceinline__ bool visible(const SceneArgs& s, f3 x, f3 n, f3 light, int32_t* lds, int32_t stride) {
    const f3 o = add(x, scale(n, kOffset));
    const f3 d = sub(light, o);
    const float dist = sqrtf(dot(d, d));
    crt_ray r;
    r.org[0] = o.x;
    r.org[1] = o.y;
    r.org[2] = o.z;
    r.tmax = dist;
    const f3 dn = scale(d, 1.0f / dist);
    r.dir[0] = dn.x;
    r.dir[1] = dn.y;
    r.dir[2] = dn.z;
    r.id = 0;
    uint32_t tests = 0;
    const crt_hit h = dev::scene_trace<FORMAT, true, 16>(s, r, 0u, kMaxIterations, tests, lds, stride);
    return !(h.t > 0.0f);
}

// Cook-Torrance terms of one light at x: f_c = A * (kd * albedo_c + spec),
// A = radiance * NL; also d(spec)/d(roughness).
struct CT {
    float A, kd, spec, dspec;
};

__device__ __forceinline__ CT cook_torrance(f3 x, f3 n, f3 eye, f3 light, float intensity, float r) {
    const f3 Lo = normalize(sub(eye, x));
    const f3 Li = neg(normalize(sub(x, light)));
    const f3 dl = sub(light, x);
    const float radiance = intensity / dot(dl, dl);
    const f3 Lh = normalize(add(Li, Lo));
    const float cosLo = fmaxf(0.0f, dot(n, Lo));
    const float cosLi = fmaxf(0.0f, dot(n, Li));
    const float dh = dot(n, Lh);
    const float cosLh = fmaxf(0.0f, dh);
    const float c = 1.0f - fmaxf(0.0f, dot(Lh, Lo));
    const float c5 = c * c * c * c * c;
    const float F = 0.04f + (1.0f - 0.04f) * c5;
    const float alpha = r * r;
    const float a2 = alpha * alpha;
    const float cl2 = cosLh * cosLh;
    // The NDF denominator cos^2 (a2 - 1) + 1 equals a2 cos^2 + sin^2.  Near a
    // highlight 1 - cos^2 cancels in FP32 (relative error ~1e-7 / sin^2, i.e.
    // 1e-3 at roughness 0.05), while |n x Lh|^2 gives sin^2 to rounding.  With
    // the half vector below the surface cos is clamped to 0 and den is 1.
    const f3 nh = cross(n, Lh);
    const float s2 = dh > 0.0f ? dot(nh, nh) : 1.0f;
    const float den = a2 * cl2 + s2;
    const float D = a2 / (kPi * den * den);
    const float dD_da2 = (s2 - a2 * cl2) / (kPi * den * den * den);   // (den - 2 a2 cos^2) / (pi den^3)
    const float dD = dD_da2 * 4.0f * r * r * r;
    const float k = (r + 1.0f) * (r + 1.0f) / 8.0f;
    const float dk = (r + 1.0f) / 4.0f;
    const float di = cosLi * (1.0f - k) + k;
    const float dv = cosLo * (1.0f - k) + k;
    const float Si = cosLi / di;
    const float So = cosLo / dv;
    const float dSi = -cosLi * (1.0f - cosLi) / (di * di);
    const float dSo = -cosLo * (1.0f - cosLo) / (dv * dv);
    const float G = Si * So;
    const float dG = (dSi * So + Si * dSo) * dk;
    const float q = fmaxf(1e-5f, 4.0f * cosLi * cosLo);
    CT t;
    t.A = radiance * cosLi;
    t.kd = 1.0f - F;
    t.spec = F * D * G / q;
    t.dspec = F * (dD * G + D * dG) / q;
    return t;
}

// Lambert term of one light at x: albedo * B, B = radiance * clamp(dot(n, Li), 0, 1).
__device__ __forceinline__ float lambert(f3 x, f3 n, f3 light, float intensity) {
    const f3 dl = sub(light, x);
    const float radiance = intensity / dot(dl, dl);
    const f3 Li = neg(normalize(sub(x, light)));
    return radiance * fminf(fmaxf(dot(n, Li), 0.0f), 1.0f);
}

__device__ __forceinline__ f3 load3(const float* p, int32_t i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

template <int FORMAT>
__global__ void __launch_bounds__(128) forward_kernel(SceneArgs s, const crt_ray* __restrict__ rays,
                                                      const crt_hit* __restrict__ hits, int64_t npix, Params P,
                                                      uint32_t seed, float* __restrict__ color,
                                                      crt_shade_record* __restrict__ rec) {
    __shared__ int32_t lds[16 * 128];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    int32_t* my = lds + threadIdx.x;
    const crt_ray r = rays[i];
    const crt_hit h = hits[i];
    crt_shade_record out;
    out.pad0 = out.pad1 = 0.0f;
    out.vis = 0;
    out.tri1 = -1;
    const f3 eye = mk(r.org[0], r.org[1], r.org[2]);
    const f3 dir = mk(r.dir[0], r.dir[1], r.dir[2]);
    out.eye[0] = eye.x;
    out.eye[1] = eye.y;
    out.eye[2] = eye.z;
    if (!(h.t > 0.0f) || h.tri < 0) {
        out.tri0 = -1;
        for (int k = 0; k < 3; ++k) out.p[k] = out.n[k] = out.q[k] = out.nq[k] = 0.0f;
        rec[i] = out;
        color[3 * i] = color[3 * i + 1] = color[3 * i + 2] = 0.0f;
        return;
    }
    const int32_t t0 = h.tri;
    const f3 p = add(eye, scale(dir, h.t));
    const f3 n = face_normal(s, t0, dir);
    // cosine-weighted bounce (Sampling.glsl:1-12) from per-pixel PCG uniforms
    const uint32_t h1 = pcg(uint32_t(i) ^ (seed * 0x9E3779B9u));
    const uint32_t h2 = pcg(h1);
    const float u1 = float(h1 >> 8) * (1.0f / 16777216.0f);
    const float u2 = float(h2 >> 8) * (1.0f / 16777216.0f);
    f3 uu = cross(n, mk(0.0f, 1.0f, 1.0f));
    if (dot(uu, uu) < 1e-12f) uu = cross(n, mk(1.0f, 0.0f, 0.0f));
    uu = normalize(uu);
    const f3 vv = cross(uu, n);
    const float ra = sqrtf(u2);
    const float rx = ra * cosf(kTwoPi * u1), ry = ra * sinf(kTwoPi * u1), rz = sqrtf(1.0f - u2);
    const f3 bd = normalize(add(add(scale(uu, rx), scale(vv, ry)), scale(n, rz)));
    const f3 bo = add(p, scale(n, kOffset));
    crt_ray br;
    br.org[0] = bo.x;
    br.org[1] = bo.y;
    br.org[2] = bo.z;
    br.tmax = 1000000.0f;
    br.dir[0] = bd.x;
    br.dir[1] = bd.y;
    br.dir[2] = bd.z;
    br.id = 0;
    uint32_t tests = 0;
    const crt_hit bh = dev::scene_trace<FORMAT, false, 16>(s, br, 0u, kMaxIterations, tests, my, 128);
    int32_t t1 = -1;
    f3 q = mk(0.0f, 0.0f, 0.0f), nq = q;
    if (bh.t > 0.0f && bh.tri >= 0) {
        t1 = bh.tri;
        q = add(bo, scale(bd, bh.t));
        nq = face_normal(s, t1, bd);
    }
    const f3 alb0 = load3(P.albedo, t0);
    const float r0 = P.rough[t0];
    const f3 alb1 = t1 >= 0 ? load3(P.albedo, t1) : mk(0.0f, 0.0f, 0.0f);
    f3 c = mk(0.0f, 0.0f, 0.0f);
    float B1 = 0.0f;
    uint32_t vis = 0;
    for (int32_t l = 0; l < P.nl; ++l) {
        const f3 L = load3(P.lpos, l);
        const float I = P.inten[l];
        if (visible<FORMAT>(s, p, n, L, my, 128)) {
            vis |= 1u << l;
            const CT t = cook_torrance(p, n, eye, L, I, r0);
            c = add(c, mk(t.A * (t.kd * alb0.x + t.spec), t.A * (t.kd * alb0.y + t.spec),
                          t.A * (t.kd * alb0.z + t.spec)));
        }
        if (t1 >= 0 && visible<FORMAT>(s, q, nq, L, my, 128)) {
            vis |= 1u << (16 + l);
            B1 += lambert(q, nq, L, I);
        }
    }
    c = add(c, mk(alb0.x * alb1.x * B1, alb0.y * alb1.y * B1, alb0.z * alb1.z * B1));
    color[3 * i] = c.x;
    color[3 * i + 1] = c.y;
    color[3 * i + 2] = c.z;
    out.tri0 = t0;
    out.tri1 = t1;
    out.vis = vis;
    out.p[0] = p.x;
    out.p[1] = p.y;
    out.p[2] = p.z;
    out.n[0] = n.x;
    out.n[1] = n.y;
    out.n[2] = n.z;
    out.q[0] = q.x;
    out.q[1] = q.y;
    out.q[2] = q.z;
    out.nq[0] = nq.x;
    out.nq[1] = nq.y;
    out.nq[2] = nq.z;
    rec[i] = out;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Adds v[0..NV) into base[idx * NV + k] for every lane with idx >= 0; lanes
// that share an index are summed in-wave first.  Called by all lanes of a wave.
template <int NV>
__device__ __forceinline__ void wave_scatter_add(float* __restrict__ base, int32_t idx, const float (&v)[NV]) {
    unsigned long long pending = __ballot(idx >= 0);
    const unsigned lane = __lane_id();
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int32_t lidx = __shfl(idx, leader);
        const bool mine = idx == lidx && idx >= 0;
        const unsigned long long m = __ballot(mine);
        if (m == (1ull << leader)) {
            // the leader alone holds this index: its own values, no reduction
            if (lane == unsigned(leader)) {
#pragma unroll
                for (int k = 0; k < NV; ++k) atomicAdd(base + int64_t(lidx) * NV + k, v[k]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                const float sum = wave_sum(mine ? v[k] : 0.0f);
                if (lane == unsigned(leader)) atomicAdd(base + int64_t(lidx) * NV + k, sum);
            }
        }
        pending &= ~m;
    }
}

// The adjoint of one pixel (a tile of 128 pixels per workgroup: every lane
// of a wave calls it, with i >= npix for lanes past the end): albedo and
// roughness gradients are added to the triangles' entries, and each light's
// intensity gradient, summed over the wave, into the workgroup's LDS sum.
template <int FORMAT>
__device__ __forceinline__ void backward_pixel(const SceneArgs& s, const crt_shade_record* __restrict__ rec,
                                               int64_t i, int64_t npix, const Params& P,
                                               const float* __restrict__ color, const float* __restrict__ target,
                                               float* __restrict__ g_albedo, float* __restrict__ g_rough,
                                               float* inten_sum, int32_t* my) {
    crt_shade_record R;
    R.tri0 = -1;
    R.tri1 = -1;
    if (i < npix) R = rec[i];
    const bool valid = R.tri0 >= 0;
    const int32_t t0 = R.tri0;
    const int32_t t1 = valid ? R.tri1 : -1;
    const f3 p = mk(R.p[0], R.p[1], R.p[2]), n = mk(R.n[0], R.n[1], R.n[2]);
    const f3 eye = mk(R.eye[0], R.eye[1], R.eye[2]);
    const f3 q = mk(R.q[0], R.q[1], R.q[2]), nq = mk(R.nq[0], R.nq[1], R.nq[2]);
    f3 g = mk(0.0f, 0.0f, 0.0f), alb0 = g, alb1 = g;
    float r0 = 0.0f;
    if (valid) {
        g = mk(2.0f * (color[3 * i] - target[3 * i]), 2.0f * (color[3 * i + 1] - target[3 * i + 1]),
               2.0f * (color[3 * i + 2] - target[3 * i + 2]));
        alb0 = load3(P.albedo, t0);
        r0 = P.rough[t0];
        if (t1 >= 0) alb1 = load3(P.albedo, t1);
    }
    float SAkd0 = 0.0f, Sd0 = 0.0f, B1 = 0.0f;
    // the light loop is wave-uniform so each light's intensity gradient is
    // reduced in-wave and added to the workgroup's sum with one LDS atomic
    for (int32_t l = 0; l < P.nl; ++l) {
        float dI = 0.0f;
        if (valid) {
            const f3 L = load3(P.lpos, l);
            const float I = P.inten[l];
            if (visible<FORMAT>(s, p, n, L, my, 128)) {
                const CT t = cook_torrance(p, n, eye, L, I, r0);
                SAkd0 += t.A * t.kd;
                Sd0 += t.A * t.dspec;
                dI += (g.x * t.A * (t.kd * alb0.x + t.spec) + g.y * t.A * (t.kd * alb0.y + t.spec) +
                       g.z * t.A * (t.kd * alb0.z + t.spec)) / I;
            }
            if (t1 >= 0 && visible<FORMAT>(s, q, nq, L, my, 128)) {
                const float b = lambert(q, nq, L, I);
                B1 += b;
                dI += (g.x * alb0.x * alb1.x + g.y * alb0.y * alb1.y + g.z * alb0.z * alb1.z) * b / I;
            }
        }
        const float sum = wave_sum(dI);
        if (__lane_id() == 0) atomicAdd(inten_sum + l, sum);
    }
    const float a0[3] = {g.x * (SAkd0 + alb1.x * B1), g.y * (SAkd0 + alb1.y * B1), g.z * (SAkd0 + alb1.z * B1)};
    const float r0g[1] = {(g.x + g.y + g.z) * Sd0};
    const float a1[3] = {g.x * alb0.x * B1, g.y * alb0.y * B1, g.z * alb0.z * B1};
    wave_scatter_add<3>(g_albedo, valid ? t0 : -1, a0);
    wave_scatter_add<1>(g_rough, valid ? t0 : -1, r0g);
    wave_scatter_add<3>(g_albedo, t1, a1);
}

// Workgroups of 128 lanes take tiles of 128 pixels in a grid-stride loop
// (the launch caps the grid at a few workgroups per CU), so the intensity
// gradients, which every pixel adds to, leave each workgroup as one global
// atomic per light instead of one per wave: the few addresses they share are
// otherwise the kernel's bottleneck.
template <int FORMAT>
__global__ void __launch_bounds__(128) backward_kernel(SceneArgs s, const crt_shade_record* __restrict__ rec,
                                                       int64_t npix, Params P, const float* __restrict__ color,
                                                       const float* __restrict__ target, float* __restrict__ g_albedo,
                                                       float* __restrict__ g_rough, float* __restrict__ g_inten) {
    __shared__ int32_t lds[16 * 128];
    __shared__ float inten_sum[kMaxLights];
    if (threadIdx.x < unsigned(kMaxLights)) inten_sum[threadIdx.x] = 0.0f;
    __syncthreads();
    int32_t* my = lds + threadIdx.x;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t base = int64_t(blockIdx.x) * blockDim.x; base < npix; base += stride)
        backward_pixel<FORMAT>(s, rec, base + threadIdx.x, npix, P, color, target, g_albedo, g_rough, inten_sum, my);
    __syncthreads();
    if (threadIdx.x < unsigned(P.nl)) atomicAdd(g_inten + threadIdx.x, inten_sum[threadIdx.x]);
}

}  // namespace shade
}  // namespace crt

namespace {

int shade_args(const crt_scene_view* scene, const crt_shade_params* params, crt::SceneArgs& a, crt::shade::Params& p) {
    if (!scene || !params) return crt::set_error(CRT_E_INVALID, "shade: null scene or params");
    if (scene->format != CRT_FORMAT_STACKLESS && scene->format != CRT_FORMAT_STACK)
        return crt::set_error(CRT_E_INVALID, "shade: unknown node format");
    if (params->nlights < 1 || params->nlights > crt::shade::kMaxLights)
        return crt::set_error(CRT_E_INVALID, "shade: nlights must be 1..16");
    if (!params->albedo || !params->roughness || !params->light_pos || !params->intensity || params->ntri <= 0)
        return crt::set_error(CRT_E_INVALID, "shade: null parameter buffer");
    a = crt::scene_args(*scene);
    p = crt::shade::Params{params->albedo, params->roughness, params->light_pos, params->intensity, params->ntri,
                           params->nlights};
    return CRT_OK;
}

}  // namespace

extern "C" {

int crt_shade_forward(const crt_scene_view* scene, const crt_ray* primary_rays, const crt_hit* primary_hits,
                      int64_t npix, const crt_shade_params* params, uint32_t seed, float* color,
                      crt_shade_record* records, void* stream) {
    crt::SceneArgs a;
    crt::shade::Params p;
    int rc = shade_args(scene, params, a, p);
    if (rc) return rc;
    if (npix < 0 || (npix > 0 && (!primary_rays || !primary_hits || !color || !records)))
        return crt::set_error(CRT_E_INVALID, "crt_shade_forward: bad buffers");
    if (npix == 0) return CRT_OK;
    const unsigned grid = unsigned((npix + 127) / 128);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (a.format == CRT_FORMAT_STACKLESS)
        hipLaunchKernelGGL(crt::shade::forward_kernel<CRT_FORMAT_STACKLESS>, dim3(grid), dim3(128), 0, st, a,
                           primary_rays, primary_hits, npix, p, seed, color, records);
    else
        hipLaunchKernelGGL(crt::shade::forward_kernel<CRT_FORMAT_STACK>, dim3(grid), dim3(128), 0, st, a,
                           primary_rays, primary_hits, npix, p, seed, color, records);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return crt::set_error(CRT_E_HIP, std::string("crt_shade_forward: ") + hipGetErrorString(e));
    return CRT_OK;
}

int crt_shade_backward(const crt_scene_view* scene, const crt_shade_record* records, int64_t npix,
                       const crt_shade_params* params, const float* color, const float* target, float* grad_albedo,
                       float* grad_roughness, float* grad_intensity, void* stream) {
    crt::SceneArgs a;
    crt::shade::Params p;
    int rc = shade_args(scene, params, a, p);
    if (rc) return rc;
    if (npix < 0 || (npix > 0 && (!records || !color || !target || !grad_albedo || !grad_roughness ||
                                  !grad_intensity)))
        return crt::set_error(CRT_E_INVALID, "crt_shade_backward: bad buffers");
    if (npix == 0) return CRT_OK;
    // at most 16 workgroups of 128 lanes per CU (more than fit at once), so
    // each workgroup walks several tiles
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    const int64_t tiles = (npix + 127) / 128;
    const unsigned grid = unsigned(tiles < int64_t(cus) * 16 ? tiles : int64_t(cus) * 16);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (a.format == CRT_FORMAT_STACKLESS)
        hipLaunchKernelGGL(crt::shade::backward_kernel<CRT_FORMAT_STACKLESS>, dim3(grid), dim3(128), 0, st, a,
                           records, npix, p, color, target, grad_albedo, grad_roughness, grad_intensity);
    else
        hipLaunchKernelGGL(crt::shade::backward_kernel<CRT_FORMAT_STACK>, dim3(grid), dim3(128), 0, st, a, records,
                           npix, p, color, target, grad_albedo, grad_roughness, grad_intensity);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return crt::set_error(CRT_E_HIP, std::string("crt_shade_backward: ") + hipGetErrorString(e));
    return CRT_OK;
}

}  // extern "C"
