/*
 * oracle_attr.c -- host restatement of GetData
 * (Source/Core/Shaders/Intersectors/Include/TraverseBVHStackless.glsl:375-408).
 *
 * TEST INFRASTRUCTURE ONLY (the checker of attr.hip).  Only tests/ load it.
 *
 * FP32, no contraction, the GLSL's order of operations:
 *   UV = uvA*u + uvB*v + uvC*w (unpackHalf2x16 of PackedData.w)
 *   N  = normalize(nA*u + nB*v + nC*w), UnpackNormal = (half x, y of
 *        PackedData.x, half x of PackedData.y), normalize = v * (1/sqrt(v.v))
 *   albedo = texture(Textures[Ref], UV).xyz if Ref > -1 and t > 0, else ModelColor.xyz
 *   emissivity, alpha = float bits of the entity's Data[0], Data[1]
 * The sampler (level 0, GL_LINEAR, GL_REPEAT, sRGB texels) is the one stated
 * in attr.hip.  Texels decode through the sRGB table of
 * candela_amd/csrc/srgb_table.h, which the tests check against the sRGB
 * formula.  Misses: normal -1, albedo 0, emissivity 0, alpha 0.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../include/candela_rt.h"
#include "../candela_amd/csrc/srgb_table.h"

static const float SRGB[256] = {CRT_SRGB_TABLE_VALUES};

/* Host view of one texture (texels in host memory). */
typedef struct {
    const uint32_t* texels;
    int32_t width, height;
} host_texture;

float oracle_half_to_float(uint32_t h) {
    uint32_t sign = (h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    if (e == 0) {
        float f = (float)m * 5.9604644775390625e-8f;
        return sign ? -f : f;
    }
    uint32_t bits = e == 31 ? (sign | 0x7F800000u | (m << 13)) : (sign | ((e + 112u) << 23) | (m << 13));
    float out;
    memcpy(&out, &bits, 4);
    return out;
}

static float wrap01(float u) {
    if (!(fabsf(u) <= 1048576.0f)) u = 0.0f;
    return u - floorf(u);
}

static void texel(const host_texture* t, int32_t i, int32_t j, float* c) {
    uint32_t v = t->texels[(int64_t)j * t->width + i];
    c[0] = SRGB[v & 0xFFu];
    c[1] = SRGB[(v >> 8) & 0xFFu];
    c[2] = SRGB[(v >> 16) & 0xFFu];
}

void oracle_sample(const uint32_t* texels, int32_t width, int32_t height, float u, float v, float* out3) {
    host_texture t = {texels, width, height};
    float x = wrap01(u) * (float)width - 0.5f;
    float y = wrap01(v) * (float)height - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float a = x - fx, b = y - fy;
    int32_t i0 = (int32_t)fx, j0 = (int32_t)fy;
    int32_t i1 = i0 + 1, j1 = j0 + 1;
    i0 = i0 < 0 ? i0 + width : i0;
    j0 = j0 < 0 ? j0 + height : j0;
    i1 = i1 >= width ? i1 - width : i1;
    j1 = j1 >= height ? j1 - height : j1;
    float t00[3], t10[3], t01[3], t11[3];
    texel(&t, i0, j0, t00);
    texel(&t, i1, j0, t10);
    texel(&t, i0, j1, t01);
    texel(&t, i1, j1, t11);
    float ia = 1.0f - a, ib = 1.0f - b;
    for (int k = 0; k < 3; ++k) out3[k] = (t00[k] * ia + t10[k] * a) * ib + (t01[k] * ia + t11[k] * a) * b;
}

/* textures: ntex host descriptors {texels (host), width, height} packed as
 * (pointer, int32 w, int32 h) like crt_texture. */
void oracle_hit_attributes(const crt_triangle* tris, const crt_vertex* verts, const crt_entity* ents,
                           const crt_texture_ref* refs, int32_t nrefs, const crt_texture* textures, int32_t ntex,
                           const crt_hit* hits, int64_t n, crt_hit_attr* out) {
    for (int64_t i = 0; i < n; ++i) {
        const crt_hit* h = &hits[i];
        crt_hit_attr* o = &out[i];
        if (h->t < 0.0f || h->mesh < 0) {
            o->normal[0] = o->normal[1] = o->normal[2] = -1.0f;
            o->emissivity = 0.0f;
            o->albedo[0] = o->albedo[1] = o->albedo[2] = 0.0f;
            o->alpha = 0.0f;
            continue;
        }
        const int32_t* tr = tris[h->tri].packed;
        const crt_vertex *A = &verts[tr[0]], *B = &verts[tr[1]], *C = &verts[tr[2]];
        float u = h->u, v = h->v, w = h->w;
        float uvx = oracle_half_to_float(A->texcoords & 0xFFFFu) * u + oracle_half_to_float(B->texcoords & 0xFFFFu) * v +
                    oracle_half_to_float(C->texcoords & 0xFFFFu) * w;
        float uvy = oracle_half_to_float(A->texcoords >> 16) * u + oracle_half_to_float(B->texcoords >> 16) * v +
                    oracle_half_to_float(C->texcoords >> 16) * w;
        float nx = oracle_half_to_float(A->normal_tangent_data[0] & 0xFFFFu) * u +
                   oracle_half_to_float(B->normal_tangent_data[0] & 0xFFFFu) * v +
                   oracle_half_to_float(C->normal_tangent_data[0] & 0xFFFFu) * w;
        float ny = oracle_half_to_float(A->normal_tangent_data[0] >> 16) * u +
                   oracle_half_to_float(B->normal_tangent_data[0] >> 16) * v +
                   oracle_half_to_float(C->normal_tangent_data[0] >> 16) * w;
        float nz = oracle_half_to_float(A->normal_tangent_data[1] & 0xFFFFu) * u +
                   oracle_half_to_float(B->normal_tangent_data[1] & 0xFFFFu) * v +
                   oracle_half_to_float(C->normal_tangent_data[1] & 0xFFFFu) * w;
        float il = 1.0f / sqrtf(nx * nx + ny * ny + nz * nz);
        float alb[3] = {0.0f, 0.0f, 0.0f};
        if (h->mesh < nrefs) {
            const crt_texture_ref* r = &refs[h->mesh];
            int textured = r->albedo > -1 && r->albedo < ntex && h->t > 0.0f;
            if (textured) {
                const crt_texture* t = &textures[r->albedo];
                textured = t->texels && t->width > 0 && t->height > 0;
                if (textured) oracle_sample(t->texels, t->width, t->height, uvx, uvy, alb);
            }
            if (!textured) memcpy(alb, r->model_color, sizeof(alb));
        }
        const crt_entity* e = &ents[h->entity];
        o->normal[0] = nx * il;
        o->normal[1] = ny * il;
        o->normal[2] = nz * il;
        memcpy(&o->emissivity, &e->data[0], 4);
        memcpy(o->albedo, alb, sizeof(alb));
        memcpy(&o->alpha, &e->data[1], 4);
    }
}

void oracle_srgb_table(float* out256) { memcpy(out256, SRGB, sizeof(SRGB)); }
