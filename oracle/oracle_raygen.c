/*
 * oracle_raygen.c -- host twin of the wavefront ray generators (raygen.hip).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/ load it.
 *
 * Restates, with the kernels' FP32 operations in the same order (no
 * contraction, no library sine / cosine):
 *   diffuse / AO  DiffuseTrace.glsl:441-446: origin = P + N * offset,
 *                 CosWeightedHemisphere(N, r) (Include/Sampling.glsl:1-12),
 *                 two uniforms for the AO ray, the next two for the GI ray;
 *   specular      SpecularTrace.glsl:509-513 and StochasticReflectionDirection
 *                 (:102-135) with SampleGGXVNDF (Include/Sampling.glsl:63-83);
 *   probes        UpdateRadianceProbes.glsl:408-423 with LambertBRDF
 *                 (:351-362) normalized.
 * P = org + dir * t and N = the hit triangle's face normal taken to world
 * space by the entity's inverse-transpose and turned against the ray.  Only
 * records with t > 0 produce rays, in input order.  The per-ray hash chain is
 * h0 = pcg(id ^ seed * 0x9E3779B9), h(k+1) = pcg(hk), uniform k = (hk >> 8) 2^-24.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../include/candela_rt.h"

void oracle_sincos_turn(float u, float* s, float* c);

typedef struct {
    float x, y, z;
} v3;

static inline v3 mk(float x, float y, float z) {
    v3 r = {x, y, z};
    return r;
}
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scale(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) { return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
static inline v3 normalize(v3 a) { return scale(a, 1.0f / sqrtf(dot(a, a))); }

static inline uint32_t pcg(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
static inline float unit(uint32_t h) { return (float)(h >> 8) * 5.9604644775390625e-8f; }

/* CosWeightedHemisphere, Include/Sampling.glsl:1-12. */
static v3 cos_weighted(v3 n, float u1, float u2) {
    float s, c;
    oracle_sincos_turn(u1, &s, &c);
    float ra = sqrtf(u2);
    float rx = ra * c, ry = ra * s, rz = sqrtf(1.0f - u2);
    v3 uu = cross(n, mk(0.0f, 1.0f, 1.0f));
    if (dot(uu, uu) < 1e-12f) uu = cross(n, mk(1.0f, 0.0f, 0.0f));
    uu = normalize(uu);
    v3 vv = cross(uu, n);
    return normalize(add(add(scale(uu, rx), scale(vv, ry)), scale(n, rz)));
}

/* SampleGGXVNDF, Include/Sampling.glsl:63-83. */
static v3 ggx_half(v3 n, float roughness, float x, float y) {
    float alpha = roughness * roughness;
    float alpha2 = alpha * alpha;
    float sp, cp;
    oracle_sincos_turn(x, &sp, &cp);
    float cos_t = sqrtf((1.0f - y) / (1.0f + (alpha2 - 1.0f) * y));
    float sin_t = sqrtf(1.0f - cos_t * cos_t);
    v3 up = fabsf(n.z) < 0.999f ? mk(0.0f, 0.0f, 1.0f) : mk(1.0f, 0.0f, 0.0f);
    v3 tangent = normalize(cross(up, n));
    v3 bitangent = cross(n, tangent);
    return normalize(add(add(scale(tangent, cp * sin_t), scale(bitangent, sp * sin_t)), scale(n, cos_t)));
}

/* GLSL reflect(I, N) = I - 2 dot(N, I) N. */
static v3 reflect3(v3 i, v3 n) { return sub(i, scale(n, 2.0f * dot(n, i))); }

static void put(crt_ray* r, v3 org, float tmax, v3 dir, uint32_t id) {
    r->org[0] = org.x;
    r->org[1] = org.y;
    r->org[2] = org.z;
    r->tmax = tmax;
    r->dir[0] = dir.x;
    r->dir[1] = dir.y;
    r->dir[2] = dir.z;
    r->id = id;
}

/* Face normal of the hit in world space, turned against d.  recs (wide
 * layouts) or tris + verts (BVH2) give the positions. */
static v3 surface(const crt_tri_record* recs, const crt_triangle* tris, const crt_vertex* verts,
                  const crt_entity* ents, const crt_hit* h, v3 d) {
    v3 a, b, c;
    if (recs) {
        const crt_tri_record* t = &recs[h->tri];
        a = mk(t->v0[0], t->v0[1], t->v0[2]);
        b = mk(t->v1[0], t->v1[1], t->v1[2]);
        c = mk(t->v2[0], t->v2[1], t->v2[2]);
    } else {
        const crt_triangle* t = &tris[h->tri];
        const float* pa = verts[t->packed[0]].position;
        const float* pb = verts[t->packed[1]].position;
        const float* pc = verts[t->packed[2]].position;
        a = mk(pa[0], pa[1], pa[2]);
        b = mk(pb[0], pb[1], pb[2]);
        c = mk(pc[0], pc[1], pc[2]);
    }
    v3 no = cross(sub(b, a), sub(c, a));
    const float* inv = ents[h->entity].inverse;
    v3 nw = normalize(mk(inv[0] * no.x + inv[1] * no.y + inv[2] * no.z, inv[4] * no.x + inv[5] * no.y + inv[6] * no.z,
                         inv[8] * no.x + inv[9] * no.y + inv[10] * no.z));
    if (dot(nw, d) > 0.0f) nw = scale(nw, -1.0f);
    return nw;
}

/* kind 0: diffuse (out0 = AO, out1 = GI); kind 1: specular (out0).  Returns
 * the number of rays written to each output.  roughness: per mesh (nmesh). */
int64_t oracle_gen_rays(int kind, const crt_tri_record* recs, const crt_triangle* tris, const crt_vertex* verts,
                        const crt_entity* ents, const crt_ray* primary, const crt_hit* hits, int64_t n,
                        const crt_raygen_params* P, const float* roughness, crt_ray* out0, crt_ray* out1) {
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i) {
        const crt_hit* h = &hits[i];
        if (!(h->t > 0.0f)) continue;
        const crt_ray* r = &primary[i];
        v3 d = mk(r->dir[0], r->dir[1], r->dir[2]);
        v3 p = add(mk(r->org[0], r->org[1], r->org[2]), scale(d, h->t));
        v3 nw = surface(recs, tris, verts, ents, h, d);
        uint32_t hs = pcg(r->id ^ (P->seed * 0x9E3779B9u));
        if (kind == 0) {
            v3 org = add(p, scale(nw, P->offset));
            float u0 = unit(hs);
            hs = pcg(hs);
            float u1 = unit(hs);
            hs = pcg(hs);
            float u2 = unit(hs);
            hs = pcg(hs);
            float u3 = unit(hs);
            if (out0) put(&out0[k], org, P->ao_tmax, cos_weighted(nw, u0, u1), r->id);
            if (out1) put(&out1[k], org, P->gi_tmax, cos_weighted(nw, u2, u3), r->id);
        } else {
            float rough = (roughness && h->mesh >= 0 && h->mesh < P->nmesh) ? roughness[h->mesh] : 1.0f;
            float t = fminf(fmaxf(rough * 1.4f, 0.0f), 1.0f);
            float off = P->spec_offset[0] * (1.0f - t) + P->spec_offset[1] * t;   /* GLSL mix */
            v3 org = add(p, scale(nw, off));
            v3 inc = normalize(d);
            float rr = rough * P->roughness_scale;
            v3 m = nw;
            if (!(rr < 0.01f)) {
                for (int tries = 0; tries < 12; ++tries) {
                    float x = unit(hs) * 0.8f;
                    hs = pcg(hs);
                    float y = unit(hs) * 0.7f;
                    hs = pcg(hs);
                    v3 hv = ggx_half(nw, rr, x, y);
                    if (dot(hv, nw) > 0.001f) {
                        m = hv;
                        break;
                    }
                }
            }
            put(&out0[k], org, P->spec_tmax, reflect3(inc, m), r->id);
        }
        ++k;
    }
    return k;
}

void oracle_gen_probe_rays(const crt_probe_grid* g, crt_ray* out) {
    int64_t rx = g->res[0], ry = g->res[1], total = rx * ry * (int64_t)g->res[2];
    for (int64_t p = 0; p < total; ++p) {
        int32_t x = (int32_t)(p % rx), y = (int32_t)((p / rx) % ry), z = (int32_t)(p / (rx * ry));
        float cx = ((float)x / (float)g->res[0]) * 2.0f - 1.0f;
        float cy = ((float)y / (float)g->res[1]) * 2.0f - 1.0f;
        float cz = ((float)z / (float)g->res[2]) * 2.0f - 1.0f;
        v3 org = mk(g->box_origin[0] + cx * g->size[0], g->box_origin[1] + cy * g->size[1],
                    g->box_origin[2] + cz * g->size[2]);
        uint32_t h = pcg((uint32_t)p ^ (g->seed * 0x9E3779B9u));
        float u0 = unit(h);
        h = pcg(h);
        float u1 = unit(h);
        float sp, cp;
        oracle_sincos_turn(u0, &sp, &cp);
        float ct = 2.0f * u1 - 1.0f;
        float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        put(&out[p], org, g->tmax, normalize(mk(st * cp, st * sp, ct)), (uint32_t)p);
    }
}
