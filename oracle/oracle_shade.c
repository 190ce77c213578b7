/*
 * oracle_shade.c -- FP64 restatement of the differentiable shading model.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of shade.hip).  Only tests/ load it.
 *
 * Given the forward pass's per-pixel records (hit points, normals, bounce hit,
 * visibility bits -- geometry and visibility held fixed), computes in FP64:
 *   color   C = sum_l CT(p, l) vis_p(l) + albedo[t0] * sum_l Lambert(q, l) vis_q(l)
 *   loss    sum over pixels and channels of (C - T)^2
 *   adjoint d(loss)/d(albedo, roughness, intensity), derived by hand
 *   FD      central finite differences of the loss, recomputing only the
 *           pixels the perturbed parameter touches
 * CT follows Source/Core/Shaders/Include/CookTorranceBRDF.glsl:1-79 with
 * metalness 0 (NDF :1-8, Schlick :10-13, GGX :15-20, FresnelSchlick :22-25,
 * CookTorranceBRDF :37-79); Lambert follows GetDirect
 * (Source/Core/Shaders/DiffuseTrace.glsl:368-372) with a point light.
 * The record layout is crt_shade_record (include/candela_rt.h).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/candela_rt.h"

#define PI_GLSL 3.14159265359

typedef struct {
    double x, y, z;
} d3;

static inline d3 mk(double x, double y, double z) {
    d3 r = {x, y, z};
    return r;
}
static inline d3 sub(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 add(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 scl(d3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 nrm(d3 a) { return scl(a, 1.0 / sqrt(dot(a, a))); }
static inline d3 f3d(const float* p) { return mk(p[0], p[1], p[2]); }
static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }

/* One light at x: f_c = A * (kd * albedo_c + spec); dspec = d(spec)/d(roughness). */
static void cook_torrance(d3 x, d3 n, d3 eye, d3 light, double intensity, double r, double* A, double* kd,
                          double* spec, double* dspec) {
    d3 Lo = nrm(sub(eye, x));
    d3 Li = scl(nrm(sub(x, light)), -1.0);
    d3 dl = sub(light, x);
    double radiance = intensity / dot(dl, dl);
    d3 Lh = nrm(add(Li, Lo));
    double cosLo = dmax(0.0, dot(n, Lo));
    double cosLi = dmax(0.0, dot(n, Li));
    double cosLh = dmax(0.0, dot(n, Lh));
    double F = 0.04 + (1.0 - 0.04) * pow(1.0 - dmax(0.0, dot(Lh, Lo)), 5.0);
    double alpha = r * r, a2 = alpha * alpha;
    double den = cosLh * cosLh * (a2 - 1.0) + 1.0;
    double D = a2 / (PI_GLSL * den * den);
    double dD = (den - 2.0 * a2 * cosLh * cosLh) / (PI_GLSL * den * den * den) * 4.0 * r * r * r;
    double k = (r + 1.0) * (r + 1.0) / 8.0, dk = (r + 1.0) / 4.0;
    double di = cosLi * (1.0 - k) + k, dv = cosLo * (1.0 - k) + k;
    double Si = cosLi / di, So = cosLo / dv;
    double dSi = -cosLi * (1.0 - cosLi) / (di * di), dSo = -cosLo * (1.0 - cosLo) / (dv * dv);
    double G = Si * So, dG = (dSi * So + Si * dSo) * dk;
    double q = dmax(1e-5, 4.0 * cosLi * cosLo);
    *A = radiance * cosLi;
    *kd = 1.0 - F;
    *spec = F * D * G / q;
    *dspec = F * (dD * G + D * dG) / q;
}

static double lambert(d3 x, d3 n, d3 light, double intensity) {
    d3 dl = sub(light, x);
    double radiance = intensity / dot(dl, dl);
    d3 Li = scl(nrm(sub(x, light)), -1.0);
    return radiance * dmin(dmax(dot(n, Li), 0.0), 1.0);
}

typedef struct {
    const double* albedo;
    const double* rough;
    const double* lpos;
    const double* inten;
    int nl;
} params_t;

/* Color of one pixel, and the per-light partial sums the adjoint needs. */
static void pixel(const crt_shade_record* R, const params_t* P, double* c, double* SAkd0, double* Sd0, double* B1,
                  double* dIc /* [nl*3] or NULL */) {
    c[0] = c[1] = c[2] = 0.0;
    *SAkd0 = *Sd0 = *B1 = 0.0;
    if (R->tri0 < 0) return;
    int t0 = R->tri0, t1 = R->tri1;
    d3 p = f3d(R->p), n = f3d(R->n), eye = f3d(R->eye), q = f3d(R->q), nq = f3d(R->nq);
    const double* a0 = P->albedo + 3 * t0;
    double r0 = P->rough[t0];
    double a1[3] = {0, 0, 0};
    if (t1 >= 0) memcpy(a1, P->albedo + 3 * t1, sizeof(a1));
    for (int l = 0; l < P->nl; ++l) {
        d3 L = mk(P->lpos[3 * l], P->lpos[3 * l + 1], P->lpos[3 * l + 2]);
        double I = P->inten[l];
        double fl[3] = {0, 0, 0};
        if (R->vis & (1u << l)) {
            double A, kd, spec, dspec;
            cook_torrance(p, n, eye, L, I, r0, &A, &kd, &spec, &dspec);
            *SAkd0 += A * kd;
            *Sd0 += A * dspec;
            for (int k = 0; k < 3; ++k) fl[k] += A * (kd * a0[k] + spec);
        }
        if (t1 >= 0 && (R->vis & (1u << (16 + l)))) {
            double b = lambert(q, nq, L, I);
            *B1 += b;
            for (int k = 0; k < 3; ++k) fl[k] += a0[k] * a1[k] * b;
        }
        for (int k = 0; k < 3; ++k) {
            c[k] += fl[k];
            if (dIc) dIc[3 * l + k] = fl[k] / I;
        }
    }
}

/* color[n*3] (FP64) for every record. */
void oracle_shade_forward(int64_t n, const crt_shade_record* rec, const double* albedo, const double* rough,
                          const double* lpos, const double* inten, int nl, double* color) {
    params_t P = {albedo, rough, lpos, inten, nl};
    for (int64_t i = 0; i < n; ++i) {
        double s0, s1, s2;
        pixel(&rec[i], &P, color + 3 * i, &s0, &s1, &s2, NULL);
    }
}

/* Loss and its exact gradient (gradients are written, not added). */
double oracle_shade_adjoint(int64_t n, const crt_shade_record* rec, const double* albedo, const double* rough,
                            const double* lpos, const double* inten, int nl, const float* target, int ntri,
                            double* g_albedo, double* g_rough, double* g_inten) {
    params_t P = {albedo, rough, lpos, inten, nl};
    memset(g_albedo, 0, sizeof(double) * 3 * (size_t)ntri);
    memset(g_rough, 0, sizeof(double) * (size_t)ntri);
    memset(g_inten, 0, sizeof(double) * (size_t)nl);
    double loss = 0.0;
    double dIc[16 * 3];
    for (int64_t i = 0; i < n; ++i) {
        const crt_shade_record* R = &rec[i];
        if (R->tri0 < 0) {
            for (int k = 0; k < 3; ++k) loss += (double)target[3 * i + k] * target[3 * i + k];
            continue;
        }
        double c[3], SAkd0, Sd0, B1;
        pixel(R, &P, c, &SAkd0, &Sd0, &B1, dIc);
        double g[3];
        for (int k = 0; k < 3; ++k) {
            double d = c[k] - (double)target[3 * i + k];
            loss += d * d;
            g[k] = 2.0 * d;
        }
        int t0 = R->tri0, t1 = R->tri1;
        const double* a0 = albedo + 3 * t0;
        double a1[3] = {0, 0, 0};
        if (t1 >= 0) memcpy(a1, albedo + 3 * t1, sizeof(a1));
        for (int k = 0; k < 3; ++k) g_albedo[3 * t0 + k] += g[k] * (SAkd0 + a1[k] * B1);
        g_rough[t0] += (g[0] + g[1] + g[2]) * Sd0;
        if (t1 >= 0)
            for (int k = 0; k < 3; ++k) g_albedo[3 * t1 + k] += g[k] * a0[k] * B1;
        for (int l = 0; l < nl; ++l)
            for (int k = 0; k < 3; ++k) g_inten[l] += g[k] * dIc[3 * l + k];
    }
    return loss;
}

static double pixel_loss(const crt_shade_record* R, const params_t* P, const float* t) {
    double c[3], s0, s1, s2;
    pixel(R, P, c, &s0, &s1, &s2, NULL);
    double l = 0.0;
    for (int k = 0; k < 3; ++k) {
        double d = c[k] - (double)t[k];
        l += d * d;
    }
    return l;
}

/* Central finite differences (loss(x + h) - loss(x - h)) / 2h for m parameters.
 * kind: 0 = albedo (index = 3 * tri + channel), 1 = roughness (index = tri),
 * 2 = intensity (index = light).  Only pixels whose records reference the
 * parameter's triangle are re-evaluated (all pixels for an intensity). */
void oracle_shade_fd(int64_t n, const crt_shade_record* rec, const double* albedo, const double* rough,
                     const double* lpos, const double* inten, int nl, const float* target, int ntri, const int* kind,
                     const int* index, int m, double h, double* out) {
    double* alb = (double*)malloc(sizeof(double) * 3 * (size_t)ntri);
    double* rgh = (double*)malloc(sizeof(double) * (size_t)ntri);
    double it[16];
    memcpy(alb, albedo, sizeof(double) * 3 * (size_t)ntri);
    memcpy(rgh, rough, sizeof(double) * (size_t)ntri);
    memcpy(it, inten, sizeof(double) * (size_t)nl);
    params_t P = {alb, rgh, lpos, it, nl};
    for (int j = 0; j < m; ++j) {
        double* slot = kind[j] == 0 ? &alb[index[j]] : kind[j] == 1 ? &rgh[index[j]] : &it[index[j]];
        int tri = kind[j] == 0 ? index[j] / 3 : kind[j] == 1 ? index[j] : -1;
        double base = *slot, lp = 0.0, lm = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            const crt_shade_record* R = &rec[i];
            if (R->tri0 < 0) continue;
            if (tri >= 0 && R->tri0 != tri && R->tri1 != tri) continue;
            *slot = base + h;
            lp += pixel_loss(R, &P, target + 3 * i);
            *slot = base - h;
            lm += pixel_loss(R, &P, target + 3 * i);
            *slot = base;
        }
        out[j] = (lp - lm) / (2.0 * h);
    }
    free(alb);
    free(rgh);
}

/* One Cook-Torrance evaluation (for closed-form tests of the restatement). */
void oracle_cook_torrance(const double* x, const double* n, const double* eye, const double* light, double intensity,
                          double r, double* out4) {
    cook_torrance(mk(x[0], x[1], x[2]), mk(n[0], n[1], n[2]), mk(eye[0], eye[1], eye[2]),
                  mk(light[0], light[1], light[2]), intensity, r, &out4[0], &out4[1], &out4[2], &out4[3]);
}
