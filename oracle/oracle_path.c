/*
 * oracle_path.c -- host twin of the multi-bounce path kernel (path.hip).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/ load it.
 *
 * For each requested path index it follows the kernel's per-path sequence
 * with the same FP32 operations in the same order (no contraction): PCG
 * draws, the jittered camera ray of GetRayDirectionAt
 * (TraverseBVHStackless.glsl:143-148), closest hits over the scene with the
 * host wide walk (oracle_trace_wide), sky radiance on escape, albedo
 * throughput, Russian roulette from the second bounce on, the face normal
 * turned against the ray, and the cosine-weighted bounce direction with the
 * kernel's quarter-angle sine / cosine polynomial.  Radiance, segment,
 * node-visit and triangle-test counts must match the kernel bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../include/candela_rt.h"

int oracle_trace_wide(int32_t width, int32_t entity_count, int64_t total_nodes, const crt_wide_group* groups,
                      const crt_tri_record* records, const crt_entity* ents, int query, uint32_t flags,
                      const crt_ray* rays, crt_hit* hits, int64_t n, int threads, uint32_t* tri_tests,
                      int32_t max_iter, int32_t stack_cap);

typedef struct {
    float x, y, z;
} p3;

static inline p3 mk3(float x, float y, float z) {
    p3 r = {x, y, z};
    return r;
}
static inline p3 add3(p3 a, p3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline p3 sub3(p3 a, p3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline p3 scale3(p3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
static inline float dot3(p3 a, p3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline p3 cross3(p3 a, p3 b) { return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
static inline p3 norm3(p3 a) { return scale3(a, 1.0f / sqrtf(dot3(a, a))); }

static inline uint32_t pcg(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

static inline float draw(uint32_t* st) {
    *st = pcg(*st);
    return (float)(*st >> 8) * 5.9604644775390625e-8f;
}

void oracle_sincos_turn(float u, float* s, float* c) {
    float u4 = u * 4.0f;
    float q = floorf(u4);
    if (q > 3.0f) q = 3.0f;
    float a = (u4 - q) * 1.57079632679f - 0.785398163397f;
    float a2 = a * a;
    float sa = a * (1.0f + a2 * (-1.66666667e-1f + a2 * (8.33333333e-3f + a2 * (-1.98412698e-4f + a2 * 2.75573192e-6f))));
    float ca = 1.0f + a2 * (-0.5f + a2 * (4.16666667e-2f + a2 * (-1.38888889e-3f + a2 * (2.48015873e-5f +
                                                                                        a2 * -2.75573192e-7f))));
    float r = 0.707106781187f;
    float sb = (sa + ca) * r, cb = (ca - sa) * r;
    int qi = (int)q;
    *s = qi == 0 ? sb : qi == 1 ? cb : qi == 2 ? -sb : -cb;
    *c = qi == 0 ? cb : qi == 1 ? -sb : qi == 2 ? -cb : sb;
}

static p3 cosine_dir(p3 n, float u1, float u2) {
    float s, c;
    oracle_sincos_turn(u1, &s, &c);
    float ra = sqrtf(u2);
    float rx = ra * c, ry = ra * s, rz = sqrtf(1.0f - u2);
    p3 uu = cross3(n, mk3(0.0f, 1.0f, 1.0f));
    if (dot3(uu, uu) < 1e-12f) uu = cross3(n, mk3(1.0f, 0.0f, 0.0f));
    uu = norm3(uu);
    p3 vv = cross3(uu, n);
    return add3(add3(scale3(uu, rx), scale3(vv, ry)), scale3(n, rz));
}

static p3 xform3(const float* m, p3 p, float w) {
    return mk3(m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12] * w, m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13] * w,
               m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14] * w);
}

/* paths[n]: the path indices to follow; out[n]. */
void oracle_trace_paths(const crt_path_params* P, int32_t width_w, int64_t nnodes, const crt_wide_group* groups,
                        const crt_tri_record* recs, const crt_entity* ents, int32_t nents, const float* albedo,
                        const int64_t* paths, int64_t n, crt_path_sample* out, int32_t stack_cap) {
    for (int64_t k = 0; k < n; ++k) {
        int64_t p = paths[k];
        uint32_t rng = pcg((uint32_t)p ^ (P->seed * 0x9E3779B9u));
        int64_t pixel = p / P->spp;
        int32_t x = (int32_t)(pixel % P->width), y = (int32_t)(pixel / P->width);
        float jx = draw(&rng), jy = draw(&rng);
        float u = ((float)x + jx) / (float)P->width, v = ((float)y + jy) / (float)P->height;
        float cx = u * 2.0f - 1.0f, cy = v * 2.0f - 1.0f;
        const float* ip = P->inv_proj;
        float ex = ip[0] * cx + ip[4] * cy + ip[8] * -1.0f + ip[12] * 1.0f;
        float ey = ip[1] * cx + ip[5] * cy + ip[9] * -1.0f + ip[13] * 1.0f;
        p3 d = xform3(P->inv_view, mk3(ex, ey, -1.0f), 0.0f);
        p3 dir = scale3(d, 1.0f / sqrtf(dot3(d, d)));
        p3 org = mk3(P->inv_view[12], P->inv_view[13], P->inv_view[14]);
        p3 thr = mk3(1.0f, 1.0f, 1.0f), rad = mk3(0.0f, 0.0f, 0.0f);
        int bounce = 0, segs = 0;
        uint32_t visits = 0, tests = 0, status = 0;
        for (;;) {
            crt_ray ray;
            ray.org[0] = org.x;
            ray.org[1] = org.y;
            ray.org[2] = org.z;
            ray.tmax = P->tmax;
            ray.dir[0] = dir.x;
            ray.dir[1] = dir.y;
            ray.dir[2] = dir.z;
            ray.id = 0;
            crt_hit h;
            uint32_t tt = 0;
            oracle_trace_wide(width_w, nents, nnodes, groups, recs, ents, CRT_QUERY_CLOSEST, 0, &ray, &h, 1, 1, &tt, 0,
                              stack_cap);
            segs++;
            visits += h.status & CRT_STATUS_VISITS_MASK;
            status |= h.status & ~CRT_STATUS_VISITS_MASK;
            tests += tt;
            if (h.tri < 0) {
                rad = add3(rad, mk3(thr.x * P->sky[0], thr.y * P->sky[1], thr.z * P->sky[2]));
                break;
            }
            p3 alb = mk3(0.0f, 0.0f, 0.0f);
            if (h.mesh >= 0 && h.mesh < P->nmesh)
                alb = mk3(albedo[3 * h.mesh], albedo[3 * h.mesh + 1], albedo[3 * h.mesh + 2]);
            thr = mk3(thr.x * alb.x, thr.y * alb.y, thr.z * alb.z);
            if (bounce >= P->max_bounces) break;
            if (bounce >= 1) {
                if (!(draw(&rng) < P->rr)) break;
                thr = mk3(thr.x / P->rr, thr.y / P->rr, thr.z / P->rr);
            }
            const crt_tri_record* r = &recs[h.tri];
            p3 a = mk3(r->v0[0], r->v0[1], r->v0[2]);
            p3 no = cross3(sub3(mk3(r->v1[0], r->v1[1], r->v1[2]), a), sub3(mk3(r->v2[0], r->v2[1], r->v2[2]), a));
            const float* inv = ents[h.entity].inverse;
            p3 nw = norm3(mk3(inv[0] * no.x + inv[1] * no.y + inv[2] * no.z, inv[4] * no.x + inv[5] * no.y + inv[6] * no.z,
                              inv[8] * no.x + inv[9] * no.y + inv[10] * no.z));
            if (dot3(nw, dir) > 0.0f) nw = scale3(nw, -1.0f);
            p3 hp = add3(org, scale3(dir, h.t));
            float u1 = draw(&rng), u2 = draw(&rng);
            org = add3(hp, scale3(nw, 1e-4f));
            dir = cosine_dir(nw, u1, u2);
            bounce++;
        }
        crt_path_sample* o = &out[k];
        o->radiance[0] = rad.x;
        o->radiance[1] = rad.y;
        o->radiance[2] = rad.z;
        o->segments = segs;
        o->visits = visits;
        o->tests = tests;
        o->status = status;
        o->pad = 0;
    }
}
