// FP32 device arithmetic shared by the traversal kernels: every function is
// the operation-for-operation restatement of one GLSL function of
// Source/Core/Shaders/Intersectors/Include/TraverseBVHStackless.glsl (and its
// twin in Include/TraverseBVHStack.glsl).  Compile with FMA contraction off.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "candela_rt.h"

#pragma clang fp contract(off)

namespace crt {
namespace dm {

// An int in LDS.  Per-lane stacks address their LDS part through an
// LDS-qualified pointer: with a generic one the compiler merges the LDS and
// the scratch branch of a pop into one flat load (a select between the two
// segment bases), and a flat load waits on every outstanding memory
// operation.
using LdsInt = __attribute__((address_space(3))) int32_t;

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

// RayTriangle (:79-97): t, or -1 when the hit is outside the triangle.
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

// RayBounds, stackless include (:100-109), t_min = 0.0001.
__device__ __forceinline__ float bounds_stackless(float4 bmin, float4 bmax, f3 ro, f3 inv, float t_max) {
    float t0x = (bmin.x - ro.x) * inv.x, t0y = (bmin.y - ro.y) * inv.y, t0z = (bmin.z - ro.z) * inv.z;
    float t1x = (bmax.x - ro.x) * inv.x, t1y = (bmax.y - ro.y) * inv.y, t1z = (bmax.z - ro.z) * inv.z;
    float tmin = mx(mx(mx(mn(t0x, t1x), mn(t0y, t1y)), mn(t0z, t1z)), 0.0001f);
    float tmax = mn(mn(mx(t0x, t1x), mn(mx(t0y, t1y), mx(t0z, t1z))), t_max);
    return (tmax >= tmin) ? tmin : -1.0f;
}

// RayBounds, stack include (:107-116).
__device__ __forceinline__ float bounds_stack(float4 bmin, float4 bmax, f3 ro, f3 inv, float maxt) {
    float fx = (bmax.x - ro.x) * inv.x, fy = (bmax.y - ro.y) * inv.y, fz = (bmax.z - ro.z) * inv.z;
    float nx = (bmin.x - ro.x) * inv.x, ny = (bmin.y - ro.y) * inv.y, nz = (bmin.z - ro.z) * inv.z;
    float tmaxx = mx(fx, nx), tmaxy = mx(fy, ny), tmaxz = mx(fz, nz);
    float tminx = mn(fx, nx), tminy = mn(fy, ny), tminz = mn(fz, nz);
    float t1 = mn(mn(tmaxx, mn(tmaxy, tmaxz)), maxt);
    float t0 = mx(mx(tminx, mx(tminy, tminz)), 0.0f);
    return (t1 >= t0) ? (t0 > 0.0f ? t0 : t1) : -1.0f;
}

// ComputeBarycentrics (:156-172).
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

__device__ __forceinline__ f3 vertex_pos(const crt_vertex* __restrict__ verts, int32_t i) {
    const float4 p = *reinterpret_cast<const float4*>(verts[i].position);
    return mk(p.x, p.y, p.z);
}

__device__ __forceinline__ uint32_t pack_status(uint32_t visits, uint32_t flags) {
    return (visits > CRT_STATUS_VISITS_MASK ? CRT_STATUS_VISITS_MASK : visits) | flags;
}

}  // namespace dm
}  // namespace crt
