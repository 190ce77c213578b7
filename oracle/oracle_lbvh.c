/*
 * oracle_lbvh.c -- host twin of the device BVH builder (gpu_build.hip).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/ load it.
 *
 * The same steps as the device, with the same FP32 operations: triangle box
 * centroids ((min + max) * 0.5), their bounds, 30-bit Morton codes of the
 * centroids quantized to 1024 cells per axis, a stable sort by code, Karras's
 * radix tree ("Maximizing Parallelism in the Construction of BVHs, Octrees,
 * and k-d Trees", HPG 2012; longest common prefix with ties broken by the
 * sorted index), boxes as unions of triangle boxes, and the stack layout with
 * two-triangle subtrees written as leaves.  Every min / max is exact, so the
 * device must produce these bytes exactly.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/candela_rt.h"

typedef struct {
    float lo[3], hi[3];
} box_t;

static box_t tri_box(const crt_vertex* v, const uint32_t* idx, int64_t t) {
    box_t b;
    const float* p0 = v[idx[3 * t + 0]].position;
    const float* p1 = v[idx[3 * t + 1]].position;
    const float* p2 = v[idx[3 * t + 2]].position;
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = fminf(fminf(p0[k], p1[k]), p2[k]);
        b.hi[k] = fmaxf(fmaxf(p0[k], p1[k]), p2[k]);
    }
    return b;
}

static box_t unite(box_t a, box_t b) {
    box_t u;
    for (int k = 0; k < 3; ++k) {
        u.lo[k] = fminf(a.lo[k], b.lo[k]);
        u.hi[k] = fmaxf(a.hi[k], b.hi[k]);
    }
    return u;
}

uint32_t oracle_spread3(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

static uint32_t quantize(float c, float lo, float scale) {
    float x = ((c - lo) * scale) * 1024.0f;
    return (uint32_t)fminf(fmaxf(x, 0.0f), 1023.0f);
}

typedef struct {
    uint32_t key, val;
} kv_t;

static int cmp_kv(const void* a, const void* b) {
    const kv_t* x = (const kv_t*)a;
    const kv_t* y = (const kv_t*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->val < y->val ? -1 : (x->val > y->val ? 1 : 0);   /* a stable sort keeps input order */
}

static int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

static int delta(const kv_t* s, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    if (s[i].key == s[j].key) return 32 + clz32((uint32_t)(i ^ j));
    return clz32(s[i].key ^ s[j].key);
}

static int32_t leaf_word(int64_t start, int32_t count) { return (int32_t)(((uint32_t)start << 4) | (uint32_t)count); }

static void put(crt_bounds* o, box_t b, int32_t wmin, int32_t wmax) {
    for (int k = 0; k < 3; ++k) {
        o->min[k] = b.lo[k];
        o->max[k] = b.hi[k];
    }
    memcpy(&o->min[3], &wmin, 4);
    memcpy(&o->max[3], &wmax, 4);
}

/* nodes [n - 1], tris [n].  Returns 0, or -1 for n < 2 or no memory. */
int oracle_build_lbvh(const crt_vertex* v, const uint32_t* idx, int64_t n, const int32_t* mesh, int32_t mesh_all,
                      int32_t t_offset, crt_stack_node* nodes, crt_triangle* tris) {
    if (n < 2) return -1;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float* cen = (float*)malloc(sizeof(float) * 3 * (size_t)n);
    kv_t* s = (kv_t*)malloc(sizeof(kv_t) * (size_t)n);
    int32_t* left = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* right = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* first = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* last = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    box_t* box = (box_t*)malloc(sizeof(box_t) * (size_t)n);
    if (!cen || !s || !left || !right || !first || !last || !box) {
        free(cen), free(s), free(left), free(right), free(first), free(last), free(box);
        return -1;
    }
    for (int64_t t = 0; t < n; ++t) {
        box_t b = tri_box(v, idx, t);
        for (int k = 0; k < 3; ++k) {
            float c = (b.lo[k] + b.hi[k]) * 0.5f;
            cen[3 * t + k] = c;
            lo[k] = fminf(lo[k], c);
            hi[k] = fmaxf(hi[k], c);
        }
    }
    float scale[3];
    for (int k = 0; k < 3; ++k) {
        float ext = hi[k] - lo[k];
        scale[k] = ext > 0.0f ? 1.0f / ext : 0.0f;
    }
    for (int64_t t = 0; t < n; ++t) {
        uint32_t x = quantize(cen[3 * t], lo[0], scale[0]);
        uint32_t y = quantize(cen[3 * t + 1], lo[1], scale[1]);
        uint32_t z = quantize(cen[3 * t + 2], lo[2], scale[2]);
        s[t].key = (oracle_spread3(x) << 2) | (oracle_spread3(y) << 1) | oracle_spread3(z);
        s[t].val = (uint32_t)t;
    }
    qsort(s, (size_t)n, sizeof(kv_t), cmp_kv);
    for (int64_t i = 0; i < n - 1; ++i) {
        int d = delta(s, n, i, i + 1) - delta(s, n, i, i - 1) >= 0 ? 1 : -1;
        int dmin = delta(s, n, i, i - d);
        int64_t lmax = 2;
        while (delta(s, n, i, i + lmax * d) > dmin) lmax *= 2;
        int64_t l = 0;
        for (int64_t t = lmax / 2; t >= 1; t /= 2)
            if (delta(s, n, i, i + (l + t) * d) > dmin) l += t;
        int64_t j = i + l * d;
        int dnode = delta(s, n, i, j);
        int64_t sp = 0, t = l;
        do {
            t = (t + 1) >> 1;
            if (delta(s, n, i, i + (sp + t) * d) > dnode) sp += t;
        } while (t > 1);
        int64_t gamma = i + sp * d + (d < 0 ? -1 : 0);
        int64_t f = i < j ? i : j, la = i < j ? j : i;
        left[i] = (int32_t)(f == gamma ? (n - 1) + gamma : gamma);
        right[i] = (int32_t)(la == gamma + 1 ? (n - 1) + gamma + 1 : gamma + 1);
        first[i] = (int32_t)f;
        last[i] = (int32_t)la;
    }
    /* boxes: children cover sub-ranges, so an internal node's box is the union
       of its range's triangle boxes; build them by increasing range length */
    int64_t* order = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n - 1));
    int64_t* cnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    if (!order || !cnt) {
        free(order), free(cnt), free(cen), free(s), free(left), free(right), free(first), free(last), free(box);
        return -1;
    }
    for (int64_t i = 0; i < n - 1; ++i) cnt[last[i] - first[i] + 1]++;
    for (int64_t k = 1; k <= n; ++k) cnt[k] += cnt[k - 1];
    for (int64_t i = n - 2; i >= 0; --i) order[--cnt[last[i] - first[i] + 1]] = i;
    for (int64_t q = 0; q < n - 1; ++q) {
        int64_t i = order[q];
        box_t b[2];
        int32_t c[2] = {left[i], right[i]};
        for (int k = 0; k < 2; ++k)
            b[k] = c[k] >= n - 1 ? tri_box(v, idx, s[c[k] - (n - 1)].val) : box[c[k]];
        box[i] = unite(b[0], b[1]);
    }
    for (int64_t i = 0; i < n; ++i) {
        uint32_t t = s[i].val;
        tris[i].packed[0] = (int32_t)idx[3 * (int64_t)t];
        tris[i].packed[1] = (int32_t)idx[3 * (int64_t)t + 1];
        tris[i].packed[2] = (int32_t)idx[3 * (int64_t)t + 2];
        tris[i].packed[3] = mesh ? mesh[t] : mesh_all;
    }
    for (int64_t i = 0; i < n - 1; ++i) {
        int32_t c[2] = {left[i], right[i]};
        crt_bounds* out[2] = {&nodes[i].left, &nodes[i].right};
        for (int k = 0; k < 2; ++k) {
            int32_t cc = c[k];
            if (cc >= n - 1) {
                int32_t leaf = cc - (int32_t)(n - 1);
                put(out[k], tri_box(v, idx, s[leaf].val), leaf_word((int64_t)t_offset + leaf, 1), 0);
            } else if (last[cc] - first[cc] == 1) {
                put(out[k], box[cc], leaf_word((int64_t)t_offset + first[cc], 2), 0);
            } else {
                put(out[k], box[cc], -1, cc);
            }
        }
    }
    free(order), free(cnt), free(cen), free(s), free(left), free(right), free(first), free(last), free(box);
    return 0;
}
