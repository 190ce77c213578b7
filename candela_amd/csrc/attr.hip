// Hit attributes: GetData (Source/Core/Shaders/Intersectors/Include/
// TraverseBVHStackless.glsl:375-408) for a buffer of closest-hit records.
//
// Per hit, in the reference's order of operations (FP32, no contraction):
//   UV     = uvA * u + uvB * v + uvC * w          (unpackHalf2x16 of PackedData.w)
//   N      = normalize(nA * u + nB * v + nC * w)  (UnpackNormal: half x, y of
//            PackedData.x, half x of PackedData.y; normalize = v * (1 / sqrt(v.v)))
//   albedo = texture(Textures[ref.albedo], UV).xyz when ref.albedo > -1,
//            else ref.model_color.xyz
//   emissivity / alpha = float bits of the entity's Data[0] / Data[1]
// where (u, v, w) are the hit's barycentrics (TUVW.yzw).  The GL sampler
// the reference binds (level 0 in a compute shader, GL_LINEAR, GL_REPEAT,
// sRGB texture) is restated as: UV wrapped to [0, 1) by u - floor(u),
// texel-space x = u * W - 0.5, bilinear weights from x - floor(x), texel
// indices wrapped, each texel's RGB decoded through the sRGB table before
// filtering.
//
// Memory per hit: the 32-B hit, the 16-B triangle, the second 16 B of each of
// the three 32-B vertices, the 32-B material record, the entity's Data[0..1],
// a 16-B texture descriptor and four 4-B texels; 32 B written.  Every load
// after the hit depends on it, so the kernel is latency-bound.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"
#include "crt_internal.h"
#include "srgb_table.h"
#include "trace_launch.h"

#pragma clang fp contract(off)

namespace crt {
namespace attr {

__constant__ float kSrgb[256] = {CRT_SRGB_TABLE_VALUES};

// unpackHalf2x16 of one half: exact (every half is a float).
__device__ __forceinline__ float half_to_float(uint32_t h) {
    const uint32_t sign = (h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    if (e == 0) {
        const float f = float(m) * 5.9604644775390625e-8f;   // m * 2^-24, exact
        return sign ? -f : f;
    }
    const uint32_t bits = e == 31 ? (sign | 0x7F800000u | (m << 13)) : (sign | ((e + 112u) << 23) | (m << 13));
    return __uint_as_float(bits);
}

__device__ __forceinline__ float wrap01(float u) {
    if (!(fabsf(u) <= 1048576.0f)) u = 0.0f;   // NaN, inf and |u| > 2^20 sample at 0
    return u - floorf(u);
}

__device__ __forceinline__ float3 texel(const crt_texture& t, int32_t i, int32_t j) {
    const uint32_t c = t.texels[int64_t(j) * t.width + i];
    return make_float3(kSrgb[c & 0xFFu], kSrgb[(c >> 8) & 0xFFu], kSrgb[(c >> 16) & 0xFFu]);
}

__device__ __forceinline__ float3 sample(const crt_texture& t, float u, float v) {
    const float x = wrap01(u) * float(t.width) - 0.5f;
    const float y = wrap01(v) * float(t.height) - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    int32_t i0 = int32_t(fx), j0 = int32_t(fy);
    int32_t i1 = i0 + 1, j1 = j0 + 1;
    i0 = i0 < 0 ? i0 + t.width : i0;
    j0 = j0 < 0 ? j0 + t.height : j0;
    i1 = i1 >= t.width ? i1 - t.width : i1;
    j1 = j1 >= t.height ? j1 - t.height : j1;
    const float3 t00 = texel(t, i0, j0), t10 = texel(t, i1, j0), t01 = texel(t, i0, j1), t11 = texel(t, i1, j1);
    const float ia = 1.0f - a, ib = 1.0f - b;
    return make_float3((t00.x * ia + t10.x * a) * ib + (t01.x * ia + t11.x * a) * b,
                       (t00.y * ia + t10.y * a) * ib + (t01.y * ia + t11.y * a) * b,
                       (t00.z * ia + t10.z * a) * ib + (t01.z * ia + t11.z * a) * b);
}

__global__ void __launch_bounds__(256) attr_kernel(SceneArgs s, const crt_texture_ref* __restrict__ refs,
                                                   int32_t nrefs, const crt_texture* __restrict__ textures,
                                                   int32_t ntex, const crt_hit* __restrict__ hits, int64_t n,
                                                   crt_hit_attr* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4* hp = reinterpret_cast<const float4*>(hits + i);
    const float4 tuvw = hp[0];
    const float4 ids = hp[1];
    const int32_t tri = __float_as_int(ids.x), mesh = __float_as_int(ids.y), ent = __float_as_int(ids.z);
    float4* o = reinterpret_cast<float4*>(out + i);
    if (tuvw.x < 0.0f || mesh < 0) {
        o[0] = make_float4(-1.0f, -1.0f, -1.0f, 0.0f);
        o[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const int4 tr = *reinterpret_cast<const int4*>(s.tris[tri].packed);
    const uint4 pa = *reinterpret_cast<const uint4*>(s.verts[tr.x].normal_tangent_data);
    const uint4 pb = *reinterpret_cast<const uint4*>(s.verts[tr.y].normal_tangent_data);
    const uint4 pc = *reinterpret_cast<const uint4*>(s.verts[tr.z].normal_tangent_data);
    const float u = tuvw.y, v = tuvw.z, w = tuvw.w;
    const float uvx = half_to_float(pa.w & 0xFFFFu) * u + half_to_float(pb.w & 0xFFFFu) * v +
                      half_to_float(pc.w & 0xFFFFu) * w;
    const float uvy = half_to_float(pa.w >> 16) * u + half_to_float(pb.w >> 16) * v + half_to_float(pc.w >> 16) * w;
    const float nx = half_to_float(pa.x & 0xFFFFu) * u + half_to_float(pb.x & 0xFFFFu) * v +
                     half_to_float(pc.x & 0xFFFFu) * w;
    const float ny = half_to_float(pa.x >> 16) * u + half_to_float(pb.x >> 16) * v + half_to_float(pc.x >> 16) * w;
    const float nz = half_to_float(pa.y & 0xFFFFu) * u + half_to_float(pb.y & 0xFFFFu) * v +
                     half_to_float(pc.y & 0xFFFFu) * w;
    const float il = 1.0f / sqrtf(nx * nx + ny * ny + nz * nz);
    float3 albedo = make_float3(0.0f, 0.0f, 0.0f);
    if (mesh < nrefs) {
        const crt_texture_ref r = refs[mesh];
        bool textured = r.albedo > -1 && r.albedo < ntex && tuvw.x > 0.0f;
        crt_texture t{};
        if (textured) {
            t = textures[r.albedo];
            textured = t.texels && t.width > 0 && t.height > 0;   // an empty texture falls back as well
        }
        if (textured) {
            albedo = sample(t, uvx, uvy);
        } else {
            albedo = make_float3(r.model_color[0], r.model_color[1], r.model_color[2]);
        }
    }
    const crt_entity& e = s.entities[ent];
    o[0] = make_float4(nx * il, ny * il, nz * il, __int_as_float(e.data[0]));
    o[1] = make_float4(albedo.x, albedo.y, albedo.z, __int_as_float(e.data[1]));
}

}  // namespace attr
}  // namespace crt

extern "C" int crt_hit_attributes(const crt_scene_view* scene, const crt_texture_ref* refs, int32_t nrefs,
                                  const crt_texture* textures, int32_t ntextures, const crt_hit* hits, int64_t n,
                                  crt_hit_attr* out, void* stream) {
    if (!scene) return crt::set_error(CRT_E_INVALID, "crt_hit_attributes: scene is null");
    if (n < 0 || (n > 0 && (!hits || !out))) return crt::set_error(CRT_E_INVALID, "crt_hit_attributes: bad buffers");
    if (nrefs < 0 || (nrefs > 0 && !refs) || ntextures < 0 || (ntextures > 0 && !textures))
        return crt::set_error(CRT_E_INVALID, "crt_hit_attributes: bad material or texture table");
    if (n > 0 && (!scene->triangles || !scene->vertices || !scene->entities))
        return crt::set_error(CRT_E_INVALID, "crt_hit_attributes: the scene needs triangles, vertices and entities");
    if (n == 0) return CRT_OK;
    const crt::SceneArgs a = crt::scene_args(*scene);
    const int64_t grid = (n + 255) / 256;
    if (grid > 0x7FFFFFFF) return crt::set_error(CRT_E_INVALID, "crt_hit_attributes: too many hits");
    hipLaunchKernelGGL(crt::attr::attr_kernel, dim3(unsigned(grid)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       a, refs, nrefs, textures, ntextures, hits, n, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return crt::set_error(CRT_E_HIP, std::string("crt_hit_attributes: kernel launch: ") + hipGetErrorString(e));
    return CRT_OK;
}
