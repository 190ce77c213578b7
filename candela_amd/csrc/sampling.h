// Device sampling helpers shared by the path kernel (path.hip) and the
// wavefront ray generators (raygen.hip).  Every function is plain FP32 with no
// contraction (the including file sets `#pragma clang fp contract(off)`) and
// uses no library sine or cosine, so the host twins in oracle/ reproduce the
// results bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_math.h"

namespace crt {
namespace smp {

using dm::cross;
using dm::dot;
using dm::f3;
using dm::mk;

// 32-bit PCG hash: one LCG step and the RXS-M-XS output permutation.
__device__ __forceinline__ uint32_t pcg(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// The top 24 bits of a hash as a uniform in [0, 1).
__device__ __forceinline__ float unit(uint32_t h) { return float(h >> 8) * 5.9604644775390625e-8f; }

__device__ __forceinline__ f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 scale(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 normalize(f3 a) { return scale(a, 1.0f / sqrtf(dot(a, a))); }

// sin and cos of 2 pi u, u in [0, 1): quadrant q = floor(4u), angle b = (4u - q)
// pi/2 in [0, pi/2), a = b - pi/4 in [-pi/4, pi/4); Taylor polynomials of sin a
// (to a^9) and cos a (to a^10), error below 4e-9; sin b = (sin a + cos a) / sqrt2,
// cos b = (cos a - sin a) / sqrt2, then the exact quadrant rotation.
__device__ __forceinline__ void sincos_turn(float u, float& s, float& c) {
    const float u4 = u * 4.0f;
    float q = floorf(u4);
    if (q > 3.0f) q = 3.0f;
    const float a = (u4 - q) * 1.57079632679f - 0.785398163397f;
    const float a2 = a * a;
    const float sa = a * (1.0f + a2 * (-1.66666667e-1f + a2 * (8.33333333e-3f + a2 * (-1.98412698e-4f +
                                                                                    a2 * 2.75573192e-6f))));
    const float ca = 1.0f + a2 * (-0.5f + a2 * (4.16666667e-2f + a2 * (-1.38888889e-3f + a2 * (2.48015873e-5f +
                                                                                            a2 * -2.75573192e-7f))));
    const float r = 0.707106781187f;
    const float sb = (sa + ca) * r, cb = (ca - sa) * r;
    const int qi = int(q);
    s = qi == 0 ? sb : qi == 1 ? cb : qi == 2 ? -sb : -cb;
    c = qi == 0 ? cb : qi == 1 ? -sb : qi == 2 ? -cb : sb;
}

// CosWeightedHemisphere (Include/Sampling.glsl:1-12) before its final
// normalize: basis uu = normalize(n x (0,1,1)) (n x (1,0,0) when n is parallel
// to (0,1,1)), vv = uu x n; radius sqrt(u2), azimuth 2 pi u1, height sqrt(1 - u2).
__device__ __forceinline__ f3 cosine_dir(f3 n, float u1, float u2) {
    float s, c;
    sincos_turn(u1, s, c);
    const float ra = sqrtf(u2);
    const float rx = ra * c, ry = ra * s, rz = sqrtf(1.0f - u2);
    f3 uu = cross(n, mk(0.0f, 1.0f, 1.0f));
    if (dot(uu, uu) < 1e-12f) uu = cross(n, mk(1.0f, 0.0f, 0.0f));
    uu = normalize(uu);
    const f3 vv = cross(uu, n);
    return add(add(scale(uu, rx), scale(vv, ry)), scale(n, rz));
}

}  // namespace smp
}  // namespace crt
