/*
 * oracle_trace.c -- CPU restatement of Candela's GLSL ray traversal.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP kernels of
 * libcandela_rt are compared against.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never calls it.
 *
 * Every function follows one GLSL function of the reference, operation for
 * operation, in FP32 with no FMA contraction (build with -ffp-contract=off):
 *   ray_triangle ........... Include/TraverseBVHStackless.glsl:79-97
 *                            (identical: Include/TraverseBVHStack.glsl:87-105)
 *   bounds_stackless ....... Include/TraverseBVHStackless.glsl:100-109
 *   bounds_stack ........... Include/TraverseBVHStack.glsl:107-116
 *   barycentrics ........... Include/TraverseBVHStackless.glsl:156-172
 *   walk_stackless ......... Include/TraverseBVHStackless.glsl:175-278 (closest)
 *                            Include/TraverseBVHStackless.glsl:463-556 (any)
 *   walk_stack ............. Include/TraverseBVHStack.glsl:168-324 (closest)
 *                            Include/TraverseBVHStack.glsl:509-657 (any)
 *   scene_closest .......... Include/TraverseBVHStackless.glsl:280-366
 *   scene_any .............. Include/TraverseBVHStackless.glsl:558-579,
 *                            Include/TraverseBVHStack.glsl:659-681
 * (shader paths relative to Source/Core/Shaders/Intersectors/)
 *
 * Semantics the GLSL leaves open, fixed here and in the HIP kernels alike:
 *   - min/max of a NaN operand returns the other operand (IEEE minNum, what
 *     fminf/fmaxf and AMD's v_min_f32/v_max_f32 do); NaN arises as 0*inf when
 *     a ray origin lies on a slab plane of an axis-parallel ray;
 *   - mat4*vec4 sums the four column terms left to right;
 *   - the node-index guard breaks on index >= u_TotalNodes (the GLSL tests
 *     `> u_TotalNodes`, which lets one read past the array).
 *
 * Pins (tests/test_oracle*.py): closed-form triangles and boxes, brute force
 * over all triangles on small scenes, stack == stackless closest-t on the same
 * tree, any-hit <=> closest-hit within tmax, crafted buffers for every guard.
 *
 * walk_wide walks the wide layout (crt_wide_build), which has no reference
 * counterpart: it is the host twin of the HIP wide walk, used to check that a
 * collapsed tree gives the same nearest hit as brute force and as the BVH2
 * walks (tests/test_wide.py), and to check the GPU wide walk bit for bit.
 */
#include <math.h>
#include <stddef.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/candela_rt.h"

#define MAX_ITER_DEFAULT 1024
#define STACK_SIZE 64

typedef struct {
    float x, y, z;
} v3;

static inline v3 mk(float x, float y, float z) {
    v3 r = {x, y, z};
    return r;
}
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline float as_float(int32_t i) {
    float f;
    memcpy(&f, &i, 4);
    return f;
}
static inline int32_t as_int(float f) {
    int32_t i;
    memcpy(&i, &f, 4);
    return i;
}
static inline float mn(float a, float b) { return fminf(a, b); }
static inline float mx(float a, float b) { return fmaxf(a, b); }

/* InverseMatrix * vec4(p, w), column-major, terms summed left to right. */
static inline v3 xform(const float* m, v3 p, float w) {
    v3 r;
    r.x = m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12] * w;
    r.y = m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13] * w;
    r.z = m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14] * w;
    return r;
}

/* RayTriangle: returns t (or -1 when outside), u, v. */
void oracle_ray_triangle(const float* ro_, const float* rd_, const float* a, const float* b,
                         const float* c, float* out_tuv) {
    v3 ro = mk(ro_[0], ro_[1], ro_[2]), rd = mk(rd_[0], rd_[1], rd_[2]);
    v3 v0 = mk(a[0], a[1], a[2]), v1 = mk(b[0], b[1], b[2]), v2 = mk(c[0], c[1], c[2]);
    v3 v1v0 = sub(v1, v0);
    v3 v2v0 = sub(v2, v0);
    v3 rov0 = sub(ro, v0);
    v3 n = cross(v1v0, v2v0);
    v3 q = cross(rov0, rd);
    float d = 1.0f / dot(rd, n);
    float u = d * dot(neg(q), v2v0);
    float v = d * dot(q, v1v0);
    float t = d * dot(neg(n), rov0);
    if (u < 0.0f || v < 0.0f || (u + v) > 1.0f) t = -1.0f;
    out_tuv[0] = t;
    out_tuv[1] = u;
    out_tuv[2] = v;
}

static inline float tri_t(v3 ro, v3 rd, v3 v0, v3 v1, v3 v2) {
    v3 v1v0 = sub(v1, v0);
    v3 v2v0 = sub(v2, v0);
    v3 rov0 = sub(ro, v0);
    v3 n = cross(v1v0, v2v0);
    v3 q = cross(rov0, rd);
    float d = 1.0f / dot(rd, n);
    float u = d * dot(neg(q), v2v0);
    float v = d * dot(q, v1v0);
    float t = d * dot(neg(n), rov0);
    if (u < 0.0f || v < 0.0f || (u + v) > 1.0f) t = -1.0f;
    return t;
}

/* RayBounds of the stackless include (t_min = 0.0001). */
static inline float bounds_stackless(const float* bmin, const float* bmax, v3 ro, v3 inv, float t_min,
                                     float t_max) {
    float t0x = (bmin[0] - ro.x) * inv.x, t0y = (bmin[1] - ro.y) * inv.y, t0z = (bmin[2] - ro.z) * inv.z;
    float t1x = (bmax[0] - ro.x) * inv.x, t1y = (bmax[1] - ro.y) * inv.y, t1z = (bmax[2] - ro.z) * inv.z;
    float tmin = mx(mx(mx(mn(t0x, t1x), mn(t0y, t1y)), mn(t0z, t1z)), t_min);
    float tmax = mn(mn(mx(t0x, t1x), mn(mx(t0y, t1y), mx(t0z, t1z))), t_max);
    return (tmax >= tmin) ? tmin : -1.0f;
}

float oracle_bounds_stackless(const float* bmin, const float* bmax, const float* ro, const float* inv,
                              float t_min, float t_max) {
    return bounds_stackless(bmin, bmax, mk(ro[0], ro[1], ro[2]), mk(inv[0], inv[1], inv[2]), t_min, t_max);
}

/* RayBounds of the stack include: entry distance, or exit distance when the
 * origin is inside. */
static inline float bounds_stack(const crt_bounds* box, v3 ro, v3 inv, float maxt) {
    float fx = (box->max[0] - ro.x) * inv.x, fy = (box->max[1] - ro.y) * inv.y, fz = (box->max[2] - ro.z) * inv.z;
    float nx = (box->min[0] - ro.x) * inv.x, ny = (box->min[1] - ro.y) * inv.y, nz = (box->min[2] - ro.z) * inv.z;
    float tmaxx = mx(fx, nx), tmaxy = mx(fy, ny), tmaxz = mx(fz, nz);
    float tminx = mn(fx, nx), tminy = mn(fy, ny), tminz = mn(fz, nz);
    float t1 = mn(mn(tmaxx, mn(tmaxy, tmaxz)), maxt);
    float t0 = mx(mx(tminx, mx(tminy, tminz)), 0.0f);
    return (t1 >= t0) ? (t0 > 0.0f ? t0 : t1) : -1.0f;
}

float oracle_bounds_stack(const float* bmin, const float* bmax, const float* ro, const float* inv, float maxt) {
    crt_bounds b;
    memcpy(b.min, bmin, 12);
    memcpy(b.max, bmax, 12);
    b.min[3] = b.max[3] = 0.0f;
    return bounds_stack(&b, mk(ro[0], ro[1], ro[2]), mk(inv[0], inv[1], inv[2]), maxt);
}

/* ComputeBarycentrics: (u, v, w) of p in triangle (a, b, c). */
static inline v3 barycentrics(v3 p, v3 a, v3 b, v3 c) {
    v3 v0 = sub(b, a), v1 = sub(c, a), v2 = sub(p, a);
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

/* Exported for the closed-form tests of ComputeBarycentrics. */
void oracle_barycentrics(const float* p, const float* a, const float* b, const float* c, float* out) {
    v3 r = barycentrics(mk(p[0], p[1], p[2]), mk(a[0], a[1], a[2]), mk(b[0], b[1], b[2]), mk(c[0], c[1], c[2]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

typedef struct {
    int32_t format;
    int32_t entity_count;
    int64_t total_nodes;
    const void* nodes;
    const crt_triangle* tris;
    const crt_vertex* verts;
    const crt_entity* ents;
    const crt_tri_record* recs; /* wide formats */
    int32_t stack_cap;          /* wide formats: stack entries before STACK_OVERFLOW */
    int32_t max_iter;
    int32_t slot_order;         /* wide formats, any hit: CRT_FLAG_SLOT_ORDER */
} scene_t;

typedef struct {
    float t;
    int32_t tri;
    int32_t mesh;
    uint32_t iters;
    uint32_t status;
    uint32_t tri_tests;
} walk_t;

static inline v3 vpos(const scene_t* s, int32_t idx) {
    const float* p = s->verts[idx].position;
    return mk(p[0], p[1], p[2]);
}

/* Tests triangles [start, start+len); returns 1 when an any-hit query is done. */
static inline int test_leaf(const scene_t* s, v3 ro, v3 rd, int32_t start, int32_t len, int any, float* tmax,
                            walk_t* w) {
    for (int32_t idx = start; idx < start + len; ++idx) {
        const crt_triangle* tr = &s->tris[idx];
        v3 a = vpos(s, tr->packed[0]), b = vpos(s, tr->packed[1]), c = vpos(s, tr->packed[2]);
        float t = tri_t(ro, rd, a, b, c);
        w->tri_tests++;
        if (t > 0.0f && t < *tmax) {
            *tmax = t;
            w->t = t;
            w->mesh = tr->packed[3];
            w->tri = idx;
            if (any) return 1;
        }
    }
    return 0;
}

/* IntersectBVHStackless / IntersectBVHStacklessOcclusion */
static void walk_stackless(const scene_t* s, v3 ro, v3 rd, int32_t start, int32_t count, float tmax, int any,
                           walk_t* w) {
    const crt_stackless_node* nodes = (const crt_stackless_node*)s->nodes;
    v3 inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    int32_t it = 0;
    int32_t ptr = start;
    int stopped = 0; /* left the loop through a break, not through its condition */
    w->t = -1.0f;
    w->tri = -1;
    w->mesh = -1;
    while (ptr >= 0 && it < s->max_iter) {
        if (ptr < start || ptr > start + count || ptr < 0 || (int64_t)ptr >= s->total_nodes) {
            w->status |= CRT_STATUS_BAD_INDEX;
            stopped = 1;
            break;
        }
        it++;
        const crt_stackless_node* n = &nodes[ptr];
        float tb = bounds_stackless(n->min, n->max, ro, inv, 0.0001f, tmax);
        if (tb > 0.0f && tb < tmax) {
            int32_t packed = as_int(n->min[3]);
            if (packed != -1) {
                if (test_leaf(s, ro, rd, packed >> 4, packed & 0xF, any, &tmax, w)) {
                    w->iters += (uint32_t)it;
                    return; /* occlusion: first accepted t */
                }
                ptr = as_int(n->max[3]);
                if (ptr < 0) {
                    stopped = 1;
                    break;
                }
                ptr += start;
                continue;
            }
            ptr++;
            continue;
        }
        ptr = as_int(n->max[3]);
        if (ptr < 0) {
            stopped = 1;
            break;
        }
        ptr += start;
    }
    if (!stopped && ptr >= 0) w->status |= CRT_STATUS_ITER_CAP;
    w->iters += (uint32_t)it;
}

/* IntersectBVHStack / IntersectBVHStackOcclusion */
static void walk_stack(const scene_t* s, v3 ro, v3 rd, int32_t start, int32_t count, float tmax, int any,
                       walk_t* w) {
    const crt_stack_node* nodes = (const crt_stack_node*)s->nodes;
    v3 inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    int32_t stack[STACK_SIZE];
    int32_t sp = 0;
    int32_t it = 0;
    int32_t cur = start;
    int stopped = 0;
    w->t = -1.0f;
    w->tri = -1;
    w->mesh = -1;
    while (it < s->max_iter) {
        if (sp >= 64 || sp < 0 || cur < start || cur > start + count || cur < 0 || (int64_t)cur >= s->total_nodes) {
            w->status |= CRT_STATUS_BAD_INDEX;
            stopped = 1;
            break;
        }
        it++;
        const crt_stack_node* n = &nodes[cur];
        int32_t lp = as_int(n->left.min[3]), rp = as_int(n->right.min[3]);
        int lleaf = lp != -1, rleaf = rp != -1;
        float lt = lleaf ? -1.0f : bounds_stack(&n->left, ro, inv, tmax);
        float rt = rleaf ? -1.0f : bounds_stack(&n->right, ro, inv, tmax);
        if (lleaf && test_leaf(s, ro, rd, lp >> 4, lp & 0xF, any, &tmax, w)) {
            w->iters += (uint32_t)it;
            return;
        }
        if (rleaf && test_leaf(s, ro, rd, rp >> 4, rp & 0xF, any, &tmax, w)) {
            w->iters += (uint32_t)it;
            return;
        }
        if (lt > 0.0f && rt > 0.0f) {
            int32_t next = as_int(n->left.max[3]) + start;
            int32_t post = as_int(n->right.max[3]) + start;
            if (rt < lt) {
                int32_t tmp = next;
                next = post;
                post = tmp;
            }
            cur = next;
            if (sp >= 63) {
                w->status |= CRT_STATUS_STACK_OVERFLOW;
                stopped = 1;
                break;
            }
            stack[sp++] = post;
            continue;
        } else if (lt > 0.0f) {
            cur = as_int(n->left.max[3]) + start;
            continue;
        } else if (rt > 0.0f) {
            cur = as_int(n->right.max[3]) + start;
            continue;
        }
        if (sp <= 0) {
            stopped = 1;
            break;
        }
        cur = stack[--sp];
    }
    if (!stopped) w->status |= CRT_STATUS_ITER_CAP;
    w->iters += (uint32_t)it;
}

/* Triangle records [start, start+len) of the wide layout. */
static inline int test_records(const scene_t* s, v3 ro, v3 rd, int32_t start, int32_t len, int any, float* tmax,
                               walk_t* w) {
    for (int32_t idx = start; idx < start + len; ++idx) {
        const crt_tri_record* r = &s->recs[idx];
        float t = tri_t(ro, rd, mk(r->v0[0], r->v0[1], r->v0[2]), mk(r->v1[0], r->v1[1], r->v1[2]),
                        mk(r->v2[0], r->v2[1], r->v2[2]));
        w->tri_tests++;
        if (t > 0.0f && t < *tmax) {
            *tmax = t;
            w->t = t;
            w->tri = r->tri;
            w->mesh = r->mesh;
            if (any) return 1;
        }
    }
    return 0;
}

/* Wide walk (the host twin of libcandela_rt's wide_walk.h, same per-ray
 * sequence):
 *   pop      with no next node, pop (node, entry distance) entries until one
 *            lies within tmax * WIDE_GROW (closest hit; any hit does not cull);
 *            an empty stack ends the walk;
 *   guards   iteration cap, then the node index range;
 *   node     every non-empty child box gets the slab arithmetic of the stack
 *            include's RayBounds with the exit distance and tmax grown by
 *            WIDE_GROW = 1 + 2^-19 (a leaf whose triangle lies on the box
 *            face, a few ulps below the computed entry, is not culled by a
 *            closer-by-ulps hit); hit leaves are appended to the pending list
 *            in slot order; hit inner children are sorted by entry distance
 *            with the kernel's compare-exchange network and pushed farthest
 *            first, all but the nearest (any hit with CRT_FLAG_SLOT_ORDER:
 *            no sort, the first hit child in slot order comes first and the
 *            others are pushed last slot first);
 *   leaves   with pending leaves the first child is pushed too (on top), the
 *            pending triangles are tested in slot order and the next node is
 *            popped; otherwise the walk goes to the first child, else pops;
 *   overflow pushes that would pass stack_cap entries end the walk before any
 *            leaf of that node is tested. */
#define WIDE_GROW (1.0f + 1.0f / 524288.0f)

/* compare-exchange: swap when entry j is strictly nearer than entry i */
static inline void cx(float* t, int32_t* n, int i, int j) {
    if (t[j] < t[i]) {
        float tt = t[i];
        int32_t nn = n[i];
        t[i] = t[j];
        n[i] = n[j];
        t[j] = tt;
        n[j] = nn;
    }
}

static void walk_wide(const scene_t* s, v3 ro, v3 rd, int32_t root, int32_t count, float tmax, int any,
                      walk_t* w) {
    const int G = s->format == CRT_FORMAT_WIDE8 ? 2 : 1;
    const crt_wide_group* groups = (const crt_wide_group*)s->nodes;
    v3 inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    int32_t st_node[256];
    float st_t[256];
    int cap = s->stack_cap < 256 ? s->stack_cap : 256;
    int sp = 0, it = 0;
    int32_t cur = root;
    w->t = -1.0f;
    w->tri = -1;
    w->mesh = -1;
    for (;;) {
        if (cur < 0) {
            int found = 0;
            while (sp > 0) {
                --sp;
                if (!any && st_t[sp] > tmax * WIDE_GROW) continue;
                cur = st_node[sp];
                found = 1;
                break;
            }
            if (!found) break;
        }
        if (it >= s->max_iter) {
            w->status |= CRT_STATUS_ITER_CAP;
            break;
        }
        if (cur < root || cur >= root + count || (int64_t)cur >= s->total_nodes) {
            w->status |= CRT_STATUS_BAD_INDEX;
            break;
        }
        it++;
        float ct[8];
        int32_t cn[8];
        int32_t pend_start[8], pend_len[8];
        int npend = 0;
        const float tgrow = tmax * WIDE_GROW;
        for (int g = 0; g < G; ++g) {
            const crt_wide_group* gp = &groups[(int64_t)cur * G + g];
            for (int k = 0; k < 4; ++k) {
                float fx = (gp->hi_x[k] - ro.x) * inv.x, fy = (gp->hi_y[k] - ro.y) * inv.y,
                      fz = (gp->hi_z[k] - ro.z) * inv.z;
                float nx = (gp->lo_x[k] - ro.x) * inv.x, ny = (gp->lo_y[k] - ro.y) * inv.y,
                      nz = (gp->lo_z[k] - ro.z) * inv.z;
                float t1 = mn(mn(mx(fx, nx), mn(mx(fy, ny), mx(fz, nz))) * WIDE_GROW, tgrow);
                float t0 = mx(mx(mn(fx, nx), mx(mn(fy, ny), mn(fz, nz))), 0.0f);
                int hit = gp->count[k] >= 0 && t1 >= t0 && (t0 > 0.0f || t1 > 0.0f);
                if (hit && gp->count[k] > 0) {
                    pend_start[npend] = gp->child[k];
                    pend_len[npend] = gp->count[k];
                    npend++;
                }
                int inner = hit && gp->count[k] == 0;
                ct[4 * g + k] = inner ? t0 : INFINITY;
                cn[4 * g + k] = inner ? gp->child[k] : -1;
            }
        }
        int m = 0;
        for (int j = 0; j < 4 * G; ++j) m += cn[j] >= 0;
        int pushes = npend ? m : (m > 0 ? m - 1 : 0);
        if (sp + pushes > cap) {
            w->status |= CRT_STATUS_STACK_OVERFLOW;
            break;
        }
        int32_t next = -1;
        if (any && s->slot_order) {
            /* slot order: the first hit child comes first, the others are
               pushed last slot first */
            int first = -1;
            for (int j = 4 * G - 1; j >= 0; --j)
                if (cn[j] >= 0) first = j;
            for (int j = 4 * G - 1; j >= 0; --j) {
                if (cn[j] >= 0 && (npend || j != first)) {
                    st_node[sp] = cn[j];
                    st_t[sp] = ct[j];
                    ++sp;
                }
            }
            next = first >= 0 ? cn[first] : -1;
        } else {
            /* ascending entry distance by the kernel's compare-exchange networks */
            if (G == 1) {
                static const int net4[5][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}};
                for (int e = 0; e < 5; ++e) cx(ct, cn, net4[e][0], net4[e][1]);
            } else {
                static const int net8[19][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6},
                                                {5, 7}, {1, 2}, {5, 6}, {0, 4}, {1, 5}, {2, 6}, {3, 7},
                                                {2, 4}, {3, 5}, {1, 2}, {3, 4}, {5, 6}};
                for (int e = 0; e < 19; ++e) cx(ct, cn, net8[e][0], net8[e][1]);
            }
            int lowest = npend ? 0 : 1;
            for (int j = 4 * G - 1; j >= lowest; --j) {
                if (j < m) {
                    st_node[sp] = cn[j];
                    st_t[sp] = ct[j];
                    ++sp;
                }
            }
            next = cn[0];
        }
        if (npend) {
            for (int j = 0; j < npend; ++j) {
                if (test_records(s, ro, rd, pend_start[j], pend_len[j], any, &tmax, w)) {
                    w->iters += (uint32_t)it;
                    return;
                }
            }
            cur = -1;
        } else {
            cur = next;
        }
    }
    w->iters += (uint32_t)it;
}

static inline void walk(const scene_t* s, v3 ro, v3 rd, const crt_entity* e, float tmax, int any, walk_t* w) {
    v3 lro = xform(e->inverse, ro, 1.0f);
    v3 lrd = xform(e->inverse, rd, 0.0f);
    if (s->format == CRT_FORMAT_STACKLESS)
        walk_stackless(s, lro, lrd, e->node_offset, e->node_count, tmax, any, w);
    else if (s->format == CRT_FORMAT_STACK)
        walk_stack(s, lro, lrd, e->node_offset, e->node_count, tmax, any, w);
    else
        walk_wide(s, lro, lrd, e->node_offset, e->node_count, tmax, any, w);
}

static inline uint32_t pack_status(uint32_t visits, uint32_t flags) {
    if (visits > CRT_STATUS_VISITS_MASK) visits = CRT_STATUS_VISITS_MASK;
    return visits | flags;
}

/* IntersectScene / IntersectSceneIgnoreTransparent (closest hit). */
static void scene_closest(const scene_t* s, const crt_ray* r, uint32_t flags, crt_hit* h, uint32_t* tri_tests) {
    v3 ro = mk(r->org[0], r->org[1], r->org[2]);
    v3 rd = mk(r->dir[0], r->dir[1], r->dir[2]);
    float closest = -1.0f;
    float tmax = r->tmax;
    int32_t mesh = -1, tri = -1, ent = -1;
    walk_t w;
    memset(&w, 0, sizeof(w));
    uint32_t visits = 0, status = 0, tests = 0;
    for (int32_t i = 0; i < s->entity_count; ++i) {
        const crt_entity* e = &s->ents[i];
        if (flags & CRT_FLAG_IGNORE_TRANSPARENT) {
            float alpha = as_float(e->data[1]);
            if (alpha < 0.99f) continue;
        }
        w.iters = 0;
        w.status = 0;
        w.tri_tests = 0;
        walk(s, ro, rd, e, tmax, 0, &w);
        visits += w.iters;
        status |= w.status;
        tests += w.tri_tests;
        if (w.t > 0.0f && w.t < tmax) {
            tmax = w.t;
            closest = w.t;
            mesh = w.mesh;
            tri = w.tri;
            ent = i;
        }
    }
    h->tri = tri;
    h->mesh = mesh;
    h->entity = ent;
    h->status = pack_status(visits, status);
    if (tri_tests) *tri_tests = tests;
    int ok = closest > 0.0f && ((flags & CRT_FLAG_TRI0_IS_MISS) ? tri > 0 : tri >= 0);
    if (ok) {
        const crt_entity* e = &s->ents[ent];
        v3 lro = xform(e->inverse, ro, 1.0f);
        v3 lrd = xform(e->inverse, rd, 0.0f);
        v3 a, b, c;
        if (s->format == CRT_FORMAT_WIDE4 || s->format == CRT_FORMAT_WIDE8) {
            const crt_tri_record* r = &s->recs[tri];
            a = mk(r->v0[0], r->v0[1], r->v0[2]);
            b = mk(r->v1[0], r->v1[1], r->v1[2]);
            c = mk(r->v2[0], r->v2[1], r->v2[2]);
        } else {
            const crt_triangle* tr = &s->tris[tri];
            a = vpos(s, tr->packed[0]);
            b = vpos(s, tr->packed[1]);
            c = vpos(s, tr->packed[2]);
        }
        v3 p = mk(lro.x + lrd.x * closest, lro.y + lrd.y * closest, lro.z + lrd.z * closest);
        v3 uvw = barycentrics(p, a, b, c);
        h->t = closest;
        h->u = uvw.x;
        h->v = uvw.y;
        h->w = uvw.z;
    } else {
        h->t = h->u = h->v = h->w = -1.0f;
    }
}

/* Occlusion IntersectScene: first entity reporting a hit wins. */
static void scene_any(const scene_t* s, const crt_ray* r, uint32_t flags, crt_hit* h, uint32_t* tri_tests) {
    v3 ro = mk(r->org[0], r->org[1], r->org[2]);
    v3 rd = mk(r->dir[0], r->dir[1], r->dir[2]);
    walk_t w;
    memset(&w, 0, sizeof(w));
    uint32_t visits = 0, status = 0, tests = 0;
    h->t = h->u = h->v = h->w = -1.0f;
    h->tri = h->mesh = h->entity = -1;
    for (int32_t i = 0; i < s->entity_count; ++i) {
        const crt_entity* e = &s->ents[i];
        if (flags & CRT_FLAG_IGNORE_TRANSPARENT) {
            float alpha = as_float(e->data[1]);
            if (alpha < 0.99f) continue;
        }
        w.iters = 0;
        w.status = 0;
        w.tri_tests = 0;
        walk(s, ro, rd, e, r->tmax, 1, &w);
        visits += w.iters;
        status |= w.status;
        tests += w.tri_tests;
        if (w.t > 0.0f) {
            h->t = w.t;
            h->tri = w.tri;
            h->mesh = w.mesh;
            h->entity = i;
            break;
        }
    }
    h->status = pack_status(visits, status);
    if (tri_tests) *tri_tests = tests;
}

typedef struct {
    const scene_t* s;
    int query;
    uint32_t flags;
    const crt_ray* rays;
    crt_hit* hits;
    uint32_t* tri_tests;
    int64_t begin, end;
} job_t;

static void* run_job(void* arg) {
    job_t* j = (job_t*)arg;
    for (int64_t i = j->begin; i < j->end; ++i) {
        uint32_t* tt = j->tri_tests ? &j->tri_tests[i] : NULL;
        if (j->query == CRT_QUERY_ANY)
            scene_any(j->s, &j->rays[i], j->flags, &j->hits[i], tt);
        else
            scene_closest(j->s, &j->rays[i], j->flags, &j->hits[i], tt);
    }
    return NULL;
}

/* Traces `n` rays on the host with `threads` POSIX threads.  `tri_tests`
 * (optional) receives the triangle tests per ray; node visits are in
 * hit.status.  max_iter <= 0 selects the reference's 1024.  Returns 0. */
static int run_scene(const scene_t* sp, int query, uint32_t flags, const crt_ray* rays, crt_hit* hits, int64_t n,
                     int threads, uint32_t* tri_tests);

int oracle_trace(int32_t format, int32_t entity_count, int64_t total_nodes, const void* nodes,
                 const crt_triangle* tris, const crt_vertex* verts, const crt_entity* ents, int query,
                 uint32_t flags, const crt_ray* rays, crt_hit* hits, int64_t n, int threads, uint32_t* tri_tests,
                 int32_t max_iter) {
    scene_t s;
    memset(&s, 0, sizeof(s));
    s.format = format;
    s.entity_count = entity_count;
    s.total_nodes = total_nodes;
    s.nodes = nodes;
    s.tris = tris;
    s.verts = verts;
    s.ents = ents;
    s.max_iter = max_iter > 0 ? max_iter : MAX_ITER_DEFAULT;
    if (format != CRT_FORMAT_STACKLESS && format != CRT_FORMAT_STACK) return -1;
    return run_scene(&s, query, flags, rays, hits, n, threads, tri_tests);
}

/* The wide layout: groups = crt_wide_group[total_nodes * width / 4],
 * records, and the entity copy whose node_offset / node_count are wide
 * roots and counts.  stack_cap = the GPU walk's LDS depth + 64. */
int oracle_trace_wide(int32_t width, int32_t entity_count, int64_t total_nodes, const crt_wide_group* groups,
                      const crt_tri_record* records, const crt_entity* ents, int query, uint32_t flags,
                      const crt_ray* rays, crt_hit* hits, int64_t n, int threads, uint32_t* tri_tests,
                      int32_t max_iter, int32_t stack_cap) {
    scene_t s;
    memset(&s, 0, sizeof(s));
    if (width != 4 && width != 8) return -1;
    s.format = width == 8 ? CRT_FORMAT_WIDE8 : CRT_FORMAT_WIDE4;
    s.entity_count = entity_count;
    s.total_nodes = total_nodes;
    s.nodes = groups;
    s.recs = records;
    s.ents = ents;
    s.stack_cap = stack_cap > 0 ? stack_cap : 80;
    s.max_iter = max_iter > 0 ? max_iter : MAX_ITER_DEFAULT;
    s.slot_order = (flags & CRT_FLAG_SLOT_ORDER) != 0;
    return run_scene(&s, query, flags, rays, hits, n, threads, tri_tests);
}

static int run_scene(const scene_t* sp, int query, uint32_t flags, const crt_ray* rays, crt_hit* hits, int64_t n,
                     int threads, uint32_t* tri_tests) {
    const scene_t s = *sp;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if ((int64_t)threads > n) threads = (int)(n > 0 ? n : 1);
    pthread_t th[256];
    job_t jobs[256];
    int64_t per = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        jobs[t].s = &s;
        jobs[t].query = query;
        jobs[t].flags = flags;
        jobs[t].rays = rays;
        jobs[t].hits = hits;
        jobs[t].tri_tests = tri_tests;
        jobs[t].begin = (int64_t)t * per;
        jobs[t].end = jobs[t].begin + per < n ? jobs[t].begin + per : n;
        if (jobs[t].begin > n) jobs[t].begin = n;
    }
    for (int t = 1; t < threads; ++t) pthread_create(&th[t], NULL, run_job, &jobs[t]);
    run_job(&jobs[0]);
    for (int t = 1; t < threads; ++t) pthread_join(th[t], NULL);
    return 0;
}

/* Brute force: for every ray, the smallest t in (0, tmax) over triangles
 * [tri_begin, tri_end) of entity 0's space, with the same RayTriangle
 * arithmetic.  The first triangle in index order wins exact ties. */
int oracle_bruteforce_closest(const crt_triangle* tris, const crt_vertex* verts, const crt_entity* ent,
                              int64_t tri_begin, int64_t tri_end, const crt_ray* rays, float* out_t,
                              int32_t* out_tri, int64_t n) {
    scene_t s;
    memset(&s, 0, sizeof(s));
    s.tris = tris;
    s.verts = verts;
    for (int64_t i = 0; i < n; ++i) {
        v3 ro = mk(rays[i].org[0], rays[i].org[1], rays[i].org[2]);
        v3 rd = mk(rays[i].dir[0], rays[i].dir[1], rays[i].dir[2]);
        v3 lro = xform(ent->inverse, ro, 1.0f);
        v3 lrd = xform(ent->inverse, rd, 0.0f);
        float tmax = rays[i].tmax;
        float best = -1.0f;
        int32_t bt = -1;
        for (int64_t k = tri_begin; k < tri_end; ++k) {
            const crt_triangle* tr = &tris[k];
            float t = tri_t(lro, lrd, vpos(&s, tr->packed[0]), vpos(&s, tr->packed[1]), vpos(&s, tr->packed[2]));
            if (t > 0.0f && t < tmax) {
                tmax = t;
                best = t;
                bt = (int32_t)k;
            }
        }
        out_t[i] = best;
        out_tri[i] = bt;
    }
    return 0;
}

/* GetRayDirectionAt + main() of Intersectors/TraverseBVHStackless.glsl:143-148,
 * 371-388: pixel-corner uv, clip (uv*2-1, -1, 1) -> inv_proj -> (xy, -1, 0) ->
 * inv_view, normalised; origin = inv_view[3].xyz.  normalize(v) = v * (1 /
 * sqrt(dot(v, v))) with a correctly rounded sqrt and division. */
void oracle_primary_rays(int width, int height, const float* inv_view, const float* inv_proj, float tmax,
                         crt_ray* rays) {
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            float u = (float)x / (float)width, v = (float)y / (float)height;
            float cx = u * 2.0f - 1.0f, cy = v * 2.0f - 1.0f;
            /* inv_proj * (cx, cy, -1, 1): x, y components */
            float ex = inv_proj[0] * cx + inv_proj[4] * cy + inv_proj[8] * -1.0f + inv_proj[12] * 1.0f;
            float ey = inv_proj[1] * cx + inv_proj[5] * cy + inv_proj[9] * -1.0f + inv_proj[13] * 1.0f;
            v3 e = mk(ex, ey, -1.0f);
            v3 d = xform(inv_view, e, 0.0f);
            float len2 = dot(d, d);
            float il = 1.0f / sqrtf(len2);
            crt_ray* r = &rays[(int64_t)y * width + x];
            r->org[0] = inv_view[12];
            r->org[1] = inv_view[13];
            r->org[2] = inv_view[14];
            r->tmax = tmax;
            r->dir[0] = d.x * il;
            r->dir[1] = d.y * il;
            r->dir[2] = d.z * il;
            r->id = (uint32_t)((int64_t)y * width + x);
        }
}

/* ---------------------------------------------------------------------------
 * Host twin of the speculative wave walk (libcandela_rt's spec_walk.h, grid
 * launch: rays [64 w, 64 w + 64) share wave w).  The wave's schedule is
 * replayed exactly, so hit records, visits and tests equal the kernel's:
 *   round      up to `inner` node-step iterations; in each, every active lane
 *              whose walk goes on and whose queue has room for `width` more
 *              leaves takes one step; the iterations stop early once no lane
 *              can step; then, when some queue holds a leaf, the leaf phase;
 *              then every lane whose walk is over goes to its next entity or
 *              writes its hit;
 *   step       pop (culling by the current tmax), guards, node: hit leaves
 *              are appended to the lane's queue, every hit inner child but
 *              the nearest is pushed and the walk goes on to the nearest;
 *   leaf phase the queued leaves of all lanes in lane order, then queue
 *              order, are list entries 0, 1, ...; order = entry * 16 +
 *              triangle of its leaf; each owner keeps the accepted test
 *              (0 < t < its tmax) with the smallest (t, order) -- any hit:
 *              (order, t) -- which ends an any-hit walk; every queue empties.
 * ------------------------------------------------------------------------- */
#define SPEC_MAX_RECORD (1 << 26)
#define SPEC_END_BATCH 32

typedef struct {
    v3 ro, rd, inv;
    float tmax, ray_tmax;
    int32_t root, count, cur, sp, it, tri, mesh, slot, ent, hit_ent;
    uint32_t tests, status, visits;
    int32_t nq, queue[16];
    int active, walk_done;
    int32_t st_node[256];
    float st_t[256];
} spec_lane_t;

typedef struct {
    const scene_t* s;
    int any;
    uint32_t flags;
    int inner, qcap;
    const crt_ray* rays;
    crt_hit* hits;
    uint32_t* tri_tests;
    int64_t n, wave_begin, wave_end;
    int stackless; /* the stackless BVH2 walk (spec_bvh2.h, D == 0) instead of the wide one */
    int stack;     /* the stack-layout BVH2 walk (spec_bvh2.h, D > 0) */
} spec_job_t;

/* Both BVH2 walks read the reference's Triangle and Vertex records. */
static inline int spec_bvh2(const spec_job_t* J) { return J->stackless || J->stack; }

static int spec_begin_walk(const spec_job_t* J, spec_lane_t* L) {
    const scene_t* s = J->s;
    int32_t e = L->ent + 1;
    while (e < s->entity_count && (J->flags & CRT_FLAG_IGNORE_TRANSPARENT) && as_float(s->ents[e].data[1]) < 0.99f)
        ++e;
    L->ent = e;
    if (e >= s->entity_count) return 0;
    const crt_entity* en = &s->ents[e];
    const crt_ray* r = &J->rays[L->slot];
    L->ro = xform(en->inverse, mk(r->org[0], r->org[1], r->org[2]), 1.0f);
    L->rd = xform(en->inverse, mk(r->dir[0], r->dir[1], r->dir[2]), 0.0f);
    L->inv = mk(1.0f / L->rd.x, 1.0f / L->rd.y, 1.0f / L->rd.z);
    if (J->any) L->tmax = L->ray_tmax;
    L->root = en->node_offset;
    L->count = en->node_count;
    L->cur = L->root;
    L->sp = 0;
    L->it = 0;
    L->nq = 0;
    /* the stackless loop `while (Pointer >= 0 ...)` never runs from a negative root */
    L->walk_done = J->stackless ? L->cur < 0 : 0;
    return 1;
}

static void spec_finish(const spec_job_t* J, spec_lane_t* L) {
    const scene_t* s = J->s;
    crt_hit* h = &J->hits[L->slot];
    h->t = h->u = h->v = h->w = -1.0f;
    const float closest = L->tri >= 0 ? L->tmax : -1.0f;
    if (J->any) {
        if (L->tri >= 0) h->t = closest;
    } else {
        int ok = closest > 0.0f && ((J->flags & CRT_FLAG_TRI0_IS_MISS) ? L->tri > 0 : L->tri >= 0);
        if (ok) {
            const crt_entity* e = &s->ents[L->hit_ent];
            const crt_ray* r = &J->rays[L->slot];
            v3 lro = xform(e->inverse, mk(r->org[0], r->org[1], r->org[2]), 1.0f);
            v3 lrd = xform(e->inverse, mk(r->dir[0], r->dir[1], r->dir[2]), 0.0f);
            v3 p = mk(lro.x + lrd.x * closest, lro.y + lrd.y * closest, lro.z + lrd.z * closest);
            v3 uvw;
            if (spec_bvh2(J)) {
                const crt_triangle* tr = &s->tris[L->tri];
                uvw = barycentrics(p, vpos(s, tr->packed[0]), vpos(s, tr->packed[1]), vpos(s, tr->packed[2]));
            } else {
                const crt_tri_record* q = &s->recs[L->tri];
                uvw = barycentrics(p, mk(q->v0[0], q->v0[1], q->v0[2]), mk(q->v1[0], q->v1[1], q->v1[2]),
                                   mk(q->v2[0], q->v2[1], q->v2[2]));
            }
            h->t = closest;
            h->u = uvw.x;
            h->v = uvw.y;
            h->w = uvw.z;
        }
    }
    h->tri = L->tri;
    h->mesh = L->mesh;
    h->entity = L->tri >= 0 ? L->hit_ent : -1;
    h->status = pack_status(L->visits, L->status);
    if (J->tri_tests) J->tri_tests[L->slot] = L->tests;
}

static int spec_can_step(const spec_job_t* J, const spec_lane_t* L) {
    if (J->stackless) return L->active && !L->walk_done && L->nq < J->qcap;        /* one leaf per node */
    if (J->stack) return L->active && !L->walk_done && L->nq <= J->qcap - 2;   /* two leaf children */
    const int width = J->s->format == CRT_FORMAT_WIDE8 ? 8 : 4;
    return L->active && !L->walk_done && L->nq <= J->qcap - width;
}

/* One node of the stackless walk (walk_stackless's step), with a hit leaf
 * queued instead of tested: the lane follows the leaf's miss link at once.
 * Returns 0 when the walk is over. */
static int sls_step(const spec_job_t* J, spec_lane_t* L) {
    const scene_t* s = J->s;
    if (L->it >= s->max_iter) {
        L->status |= CRT_STATUS_ITER_CAP;
        return 0;
    }
    if (L->cur < L->root || L->cur > L->root + L->count || (int64_t)L->cur >= s->total_nodes) {
        L->status |= CRT_STATUS_BAD_INDEX;
        return 0;
    }
    L->it++;
    const crt_stackless_node* nd = &((const crt_stackless_node*)s->nodes)[L->cur];
    float tb = bounds_stackless(nd->min, nd->max, L->ro, L->inv, 0.0001f, L->tmax);
    const int32_t link = as_int(nd->max[3]);
    const int32_t packed = as_int(nd->min[3]);
    const int hit = tb > 0.0f && tb < L->tmax;
    if (hit && packed == -1) {
        L->cur++;
        return 1;
    }
    if (hit) L->queue[L->nq++] = packed;
    if (link < 0) return 0;
    L->cur = link + L->root;
    return 1;
}

/* One entry of the stack walk (walk_stack's step), with the non-empty leaf
 * children queued (left, then right) instead of tested; the next entry is
 * decided from the inner children's box distances as in walk_stack.  A push
 * at 63 entries ends the walk with STACK_OVERFLOW after the node's leaves are
 * queued.  Returns 0 when the walk is over. */
static int stk_step(const spec_job_t* J, spec_lane_t* L) {
    const scene_t* s = J->s;
    if (L->it >= s->max_iter) {
        L->status |= CRT_STATUS_ITER_CAP;
        return 0;
    }
    if (L->cur < L->root || L->cur > L->root + L->count || L->cur < 0 || (int64_t)L->cur >= s->total_nodes) {
        L->status |= CRT_STATUS_BAD_INDEX;
        return 0;
    }
    L->it++;
    const crt_stack_node* nd = &((const crt_stack_node*)s->nodes)[L->cur];
    const int32_t lp = as_int(nd->left.min[3]), rp = as_int(nd->right.min[3]);
    const float lt = lp != -1 ? -1.0f : bounds_stack(&nd->left, L->ro, L->inv, L->tmax);
    const float rt = rp != -1 ? -1.0f : bounds_stack(&nd->right, L->ro, L->inv, L->tmax);
    if (lp != -1 && (lp & 0xF) > 0) L->queue[L->nq++] = lp;
    if (rp != -1 && (rp & 0xF) > 0) L->queue[L->nq++] = rp;
    if (lt > 0.0f && rt > 0.0f) {
        int32_t next = as_int(nd->left.max[3]) + L->root;
        int32_t post = as_int(nd->right.max[3]) + L->root;
        if (rt < lt) {
            const int32_t tmp = next;
            next = post;
            post = tmp;
        }
        if (L->sp >= 63) {
            L->status |= CRT_STATUS_STACK_OVERFLOW;
            return 0;
        }
        L->st_node[L->sp++] = post;
        L->cur = next;
    } else if (lt > 0.0f) {
        L->cur = as_int(nd->left.max[3]) + L->root;
    } else if (rt > 0.0f) {
        L->cur = as_int(nd->right.max[3]) + L->root;
    } else if (L->sp <= 0) {
        return 0;
    } else {
        L->cur = L->st_node[--L->sp];
    }
    return 1;
}

static int spec_step(const spec_job_t* J, spec_lane_t* L) {
    const scene_t* s = J->s;
    const int G = s->format == CRT_FORMAT_WIDE8 ? 2 : 1;
    const crt_wide_group* groups = (const crt_wide_group*)s->nodes;
    if (L->cur < 0) {
        int found = 0;
        while (L->sp > 0) {
            --L->sp;
            if (!J->any && L->st_t[L->sp] > L->tmax * WIDE_GROW) continue;
            L->cur = L->st_node[L->sp];
            found = 1;
            break;
        }
        if (!found) return 0;
    }
    if (L->it >= s->max_iter) {
        L->status |= CRT_STATUS_ITER_CAP;
        return 0;
    }
    if (L->cur < L->root || L->cur >= L->root + L->count || (int64_t)L->cur >= s->total_nodes) {
        L->status |= CRT_STATUS_BAD_INDEX;
        return 0;
    }
    L->it++;
    float ct[8];
    int32_t cn[8];
    int32_t leaves[8];
    int nl = 0;
    const float tgrow = L->tmax * WIDE_GROW;
    for (int j = 0; j < 8; ++j) {
        ct[j] = INFINITY;
        cn[j] = -1;
    }
    for (int g = 0; g < G; ++g) {
        const crt_wide_group* gp = &groups[(int64_t)L->cur * G + g];
        for (int k = 0; k < 4; ++k) {
            v3 ro = L->ro, inv = L->inv;
            float fx = (gp->hi_x[k] - ro.x) * inv.x, fy = (gp->hi_y[k] - ro.y) * inv.y,
                  fz = (gp->hi_z[k] - ro.z) * inv.z;
            float nx = (gp->lo_x[k] - ro.x) * inv.x, ny = (gp->lo_y[k] - ro.y) * inv.y,
                  nz = (gp->lo_z[k] - ro.z) * inv.z;
            float t1 = mn(mn(mx(fx, nx), mn(mx(fy, ny), mx(fz, nz))) * WIDE_GROW, tgrow);
            float t0 = mx(mx(mn(fx, nx), mx(mn(fy, ny), mn(fz, nz))), 0.0f);
            int32_t cnt = gp->count[k], chi = gp->child[k];
            int hit = cnt >= 0 && t1 >= t0 && t1 > 0.0f;
            if (hit && cnt > 0) {
                if (chi >= 0 && chi <= SPEC_MAX_RECORD - cnt)
                    leaves[nl++] = (chi << 4) | cnt;
                else
                    L->status |= CRT_STATUS_BAD_INDEX;
            }
            int inner = hit && cnt == 0;
            ct[4 * g + k] = inner ? t0 : INFINITY;
            cn[4 * g + k] = inner ? chi : -1;
        }
    }
    int m = 0;
    for (int j = 0; j < 8; ++j) m += cn[j] >= 0;
    int pushes = m > 0 ? m - 1 : 0;
    if (L->sp + pushes > s->stack_cap) {   /* the node's leaves are dropped with the walk */
        L->status |= CRT_STATUS_STACK_OVERFLOW;
        return 0;
    }
    if (G == 1) {
        static const int net4[5][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}};
        for (int e = 0; e < 5; ++e) cx(ct, cn, net4[e][0], net4[e][1]);
    } else {
        static const int net8[19][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6},
                                        {5, 7}, {1, 2}, {5, 6}, {0, 4}, {1, 5}, {2, 6}, {3, 7},
                                        {2, 4}, {3, 5}, {1, 2}, {3, 4}, {5, 6}};
        for (int e = 0; e < 19; ++e) cx(ct, cn, net8[e][0], net8[e][1]);
    }
    for (int j = 4 * G - 1; j >= 1; --j)
        if (j < m) {
            L->st_node[L->sp] = cn[j];
            L->st_t[L->sp] = ct[j];
            L->sp++;
        }
    L->cur = cn[0];
    for (int j = 0; j < nl; ++j) L->queue[L->nq++] = leaves[j];
    return 1;
}

/* Tests the queued leaves of the wave in list order and gives each owner its
 * winner; every queue is empty afterwards. */
static void spec_leaf_phase(const spec_job_t* J, spec_lane_t* lanes, uint16_t* list) {
    const scene_t* s = J->s;
    int32_t total = 0;
    for (int l = 0; l < 64; ++l) {
        spec_lane_t* L = &lanes[l];
        for (int j = 0; j < L->nq; ++j) {
            list[total++] = (uint16_t)((l << 4) | j);
            L->tests += (uint32_t)(L->queue[j] & 0xF);
        }
    }
    uint64_t keys[64];
    for (int l = 0; l < 64; ++l) keys[l] = ~0ull;
    for (int32_t idx = 0; idx < total; ++idx) {
        const int owner = list[idx] >> 4;
        spec_lane_t* O = &lanes[owner];
        const int32_t leaf = O->queue[list[idx] & 15];
        const int32_t first = leaf >> 4, len = leaf & 0xF;
        for (int32_t x = 0; x < len; ++x) {
            float t;
            if (spec_bvh2(J)) {
                const crt_triangle* tr = &s->tris[first + x];
                t = tri_t(O->ro, O->rd, vpos(s, tr->packed[0]), vpos(s, tr->packed[1]), vpos(s, tr->packed[2]));
            } else {
                const crt_tri_record* r = &s->recs[first + x];
                t = tri_t(O->ro, O->rd, mk(r->v0[0], r->v0[1], r->v0[2]), mk(r->v1[0], r->v1[1], r->v1[2]),
                          mk(r->v2[0], r->v2[1], r->v2[2]));
            }
            if (t > 0.0f && t < O->tmax) {
                uint32_t tb;
                memcpy(&tb, &t, 4);
                const uint64_t order = ((uint64_t)(uint32_t)idx << 4) | (uint32_t)x;
                uint64_t key = J->any ? ((order << 32) | tb) : (((uint64_t)tb << 32) | order);
                if (key < keys[owner]) keys[owner] = key;
            }
        }
    }
    for (int l = 0; l < 64; ++l) {
        spec_lane_t* L = &lanes[l];
        if (keys[l] != ~0ull) {
            const uint32_t order = J->any ? (uint32_t)(keys[l] >> 32) : (uint32_t)keys[l];
            const uint32_t tb = J->any ? (uint32_t)keys[l] : (uint32_t)(keys[l] >> 32);
            const int32_t leaf = L->queue[list[order >> 4] & 15];
            const int32_t idx = (leaf >> 4) + (int32_t)(order & 15u);
            memcpy(&L->tmax, &tb, 4);
            if (spec_bvh2(J)) {
                L->tri = idx;
                L->mesh = s->tris[idx].packed[3];
            } else {
                L->tri = s->recs[idx].tri;
                L->mesh = s->recs[idx].mesh;
            }
            L->hit_ent = L->ent;
            if (J->any) L->walk_done = 1;
        }
    }
    for (int l = 0; l < 64; ++l) lanes[l].nq = 0;
}

static void* spec_run_job(void* arg) {
    const spec_job_t* J = (const spec_job_t*)arg;
    spec_lane_t* lanes = (spec_lane_t*)malloc(sizeof(spec_lane_t) * 64);
    uint16_t* list = (uint16_t*)malloc(sizeof(uint16_t) * 64 * 16);
    if (!lanes || !list) {
        free(lanes);
        free(list);
        return NULL;
    }
    for (int64_t w = J->wave_begin; w < J->wave_end; ++w) {
        for (int l = 0; l < 64; ++l) {
            spec_lane_t* L = &lanes[l];
            memset(L, 0, offsetof(spec_lane_t, st_node));
            int64_t i = w * 64 + l;
            if (i >= J->n) continue;
            L->active = 1;
            L->slot = (int32_t)i;
            L->ray_tmax = J->rays[i].tmax;
            L->tmax = L->ray_tmax;
            L->ent = -1;
            L->hit_ent = -1;
            L->tri = L->mesh = -1;
            if (!spec_begin_walk(J, L)) {
                spec_finish(J, L);
                L->active = 0;
            }
        }
        for (;;) {
            int act = 0;
            for (int l = 0; l < 64; ++l) act |= lanes[l].active;
            if (!act) break;
            for (int k = 0; k < J->inner; ++k) {
                for (int l = 0; l < 64; ++l) {
                    spec_lane_t* L = &lanes[l];
                    if (spec_can_step(J, L))
                        if (!(J->stackless ? sls_step(J, L) : J->stack ? stk_step(J, L) : spec_step(J, L)))
                            L->walk_done = 1;
                }
                int stepping = 0;
                for (int l = 0; l < 64; ++l) stepping |= spec_can_step(J, &lanes[l]);
                if (!stepping) break;
            }
            int queued = 0;
            for (int l = 0; l < 64; ++l) queued |= lanes[l].nq > 0;
            if (queued) spec_leaf_phase(J, lanes, list);
            /* the walk ends run once SPEC_END_BATCH lanes wait or no lane can
               step (spec_walk.h: kEndBatch) */
            int ending = 0, can = 0;
            for (int l = 0; l < 64; ++l) {
                ending += lanes[l].active && lanes[l].walk_done;
                can |= spec_can_step(J, &lanes[l]);
            }
            if (!ending || (ending < SPEC_END_BATCH && can)) continue;
            for (int l = 0; l < 64; ++l) {
                spec_lane_t* L = &lanes[l];
                if (L->active && L->walk_done) {
                    L->visits += (uint32_t)L->it;
                    int fin = J->any && L->tri >= 0;
                    if (!fin) fin = !spec_begin_walk(J, L);
                    if (fin) {
                        spec_finish(J, L);
                        L->active = 0;
                    }
                }
            }
        }
    }
    free(lanes);
    free(list);
    return NULL;
}

/* Schedule parameters of the kernel's preset: stack_cap = its LDS stack
 * entries + 64, inner node steps per round and the queue slots per lane
 * (width..16). */
int oracle_trace_wide_spec(int32_t width, int32_t entity_count, int64_t total_nodes, const crt_wide_group* groups,
                           const crt_tri_record* records, const crt_entity* ents, int query, uint32_t flags,
                           const crt_ray* rays, crt_hit* hits, int64_t n, int threads, uint32_t* tri_tests,
                           int32_t max_iter, int32_t stack_cap, int32_t inner, int32_t queue_slots) {
    scene_t s;
    memset(&s, 0, sizeof(s));
    if ((width != 4 && width != 8) || inner < 1 || queue_slots < width || queue_slots > 16 || stack_cap < 1 ||
        stack_cap > 248)
        return -1;
    s.format = width == 8 ? CRT_FORMAT_WIDE8 : CRT_FORMAT_WIDE4;
    s.entity_count = entity_count;
    s.total_nodes = total_nodes;
    s.nodes = groups;
    s.recs = records;
    s.ents = ents;
    s.stack_cap = stack_cap;
    s.max_iter = max_iter > 0 ? max_iter : MAX_ITER_DEFAULT;
    int64_t waves = (n + 63) / 64;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if ((int64_t)threads > waves) threads = (int)(waves > 0 ? waves : 1);
    pthread_t th[256];
    spec_job_t jobs[256];
    int64_t per = (waves + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        spec_job_t* J = &jobs[t];
        J->s = &s;
        J->any = query == CRT_QUERY_ANY;
        J->flags = flags;
        J->inner = inner;
        J->qcap = queue_slots;
        J->rays = rays;
        J->hits = hits;
        J->tri_tests = tri_tests;
        J->n = n;
        J->wave_begin = (int64_t)t * per < waves ? (int64_t)t * per : waves;
        J->wave_end = J->wave_begin + per < waves ? J->wave_begin + per : waves;
        J->stackless = 0;
        J->stack = 0;
    }
    for (int t = 1; t < threads; ++t) pthread_create(&th[t], NULL, spec_run_job, &jobs[t]);
    spec_run_job(&jobs[0]);
    for (int t = 1; t < threads; ++t) pthread_join(th[t], NULL);
    return 0;
}

/* Runs the BVH2 twin: stack = 0 for the stackless layout, 1 for the stack
 * layout. */
static int spec_bvh2_run(int stack, int32_t entity_count, int64_t total_nodes, const void* nodes,
                         const crt_triangle* tris, const crt_vertex* verts, const crt_entity* ents, int query,
                         uint32_t flags, const crt_ray* rays, crt_hit* hits, int64_t n, int threads,
                         uint32_t* tri_tests, int32_t max_iter, int32_t inner, int32_t queue_slots) {
    scene_t s;
    memset(&s, 0, sizeof(s));
    if (inner < 1 || queue_slots < (stack ? 2 : 1) || queue_slots > 16) return -1;
    s.format = stack ? CRT_FORMAT_STACK : CRT_FORMAT_STACKLESS;
    s.entity_count = entity_count;
    s.total_nodes = total_nodes;
    s.nodes = nodes;
    s.tris = tris;
    s.verts = verts;
    s.ents = ents;
    s.max_iter = max_iter > 0 ? max_iter : MAX_ITER_DEFAULT;
    int64_t waves = (n + 63) / 64;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if ((int64_t)threads > waves) threads = (int)(waves > 0 ? waves : 1);
    pthread_t th[256];
    spec_job_t jobs[256];
    int64_t per = (waves + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        spec_job_t* J = &jobs[t];
        J->s = &s;
        J->any = query == CRT_QUERY_ANY;
        J->flags = flags;
        J->inner = inner;
        J->qcap = queue_slots;
        J->rays = rays;
        J->hits = hits;
        J->tri_tests = tri_tests;
        J->n = n;
        J->wave_begin = (int64_t)t * per < waves ? (int64_t)t * per : waves;
        J->wave_end = J->wave_begin + per < waves ? J->wave_begin + per : waves;
        J->stackless = !stack;
        J->stack = stack;
    }
    for (int t = 1; t < threads; ++t) pthread_create(&th[t], NULL, spec_run_job, &jobs[t]);
    spec_run_job(&jobs[0]);
    for (int t = 1; t < threads; ++t) pthread_join(th[t], NULL);
    return 0;
}

/* The host twin of the speculative stackless walk's grid launch
 * (spec_bvh2.h, D == 0): 64 consecutive rays per wave, `inner` node steps per
 * round, `queue_slots` queued leaves per lane (1..16). */
int oracle_trace_stackless_spec(int32_t entity_count, int64_t total_nodes, const crt_stackless_node* nodes,
                                const crt_triangle* tris, const crt_vertex* verts, const crt_entity* ents, int query,
                                uint32_t flags, const crt_ray* rays, crt_hit* hits, int64_t n, int threads,
                                uint32_t* tri_tests, int32_t max_iter, int32_t inner, int32_t queue_slots) {
    return spec_bvh2_run(0, entity_count, total_nodes, nodes, tris, verts, ents, query, flags, rays, hits, n, threads,
                         tri_tests, max_iter, inner, queue_slots);
}

/* The host twin of the speculative stack-layout walk's grid launch
 * (spec_bvh2.h, D > 0): as above with `queue_slots` in 2..16 (a node queues
 * up to two leaf children). */
int oracle_trace_stack_spec(int32_t entity_count, int64_t total_nodes, const crt_stack_node* nodes,
                            const crt_triangle* tris, const crt_vertex* verts, const crt_entity* ents, int query,
                            uint32_t flags, const crt_ray* rays, crt_hit* hits, int64_t n, int threads,
                            uint32_t* tri_tests, int32_t max_iter, int32_t inner, int32_t queue_slots) {
    return spec_bvh2_run(1, entity_count, total_nodes, nodes, tris, verts, ents, query, flags, rays, hits, n, threads,
                         tri_tests, max_iter, inner, queue_slots);
}
