/*
 * mesh_hits.c -- CPU counterpart of the projector's mesh pass, per ray.
 *
 * TEST INFRASTRUCTURE ONLY (see transcription.c).  It produces, for one ray,
 * what deepdrr obtains from OpenGL and its peel post-processing kernels and
 * then hands to projectKernel:
 *   - per layer, the subtractive hit list (alpha, facing) after
 *     peel_postprocess_kernel.cu:15-158 `tide` (cutoff, sort, duplicate
 *     removal, gap filling, "sea level" filtering, gap filling), restated
 *     statement by statement;
 *   - per (layer, unique material), the additive sum of the density shader
 *     (pyrenderdrr/shaders/density.frag: distance x facing x density, summed
 *     by additive blending, projector.py:1063-1088) with the overlaps of
 *     higher subtractive layers taken off as density_between.frag does
 *     (projector.py:1258-1306), and the facing sum (the second channel).
 * The rasterised fragments of the reference are replaced by exact
 * ray-triangle intersections (Moller-Trumbore, brute force over every
 * triangle, FP32 in the operation order written below), sorted by
 * (alpha, triangle index).  Entering a mesh (counter-clockwise winding seen
 * from outside) has subtractive facing +1 and additive facing -1: the z-flip
 * of the reference's OpenGL camera (projector.py:866-875) makes the faces a
 * ray enters back-facing, so density.frag's sum is +density x chord length.
 * Deviations, the same in the HIP kernels: a hit repeating the previous kept
 * hit of its list with the same facing within 1e-4 mm + 1e-6 alpha is a
 * duplicate (a ray through an edge shared by two triangles; rasterisation
 * counts such a point once), and at most DRR_MESH_RAY_HITS hits per ray and
 * max_hits per layer, the nearest, are kept.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mesh_hits.h"

/* Moller-Trumbore in FP32, operation order fixed (no contraction). */
int ora_ray_triangle(const float o[3], const float d[3], const float v0[3], const float v1[3],
                     const float v2[3], float *t_out, int *entry) {
    float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    float pv[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    float det = (e1[0] * pv[0] + e1[1] * pv[1]) + e1[2] * pv[2];
    if (det == 0.0f) return 0;
    float inv = 1.0f / det;
    float tv[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    float u = ((tv[0] * pv[0] + tv[1] * pv[1]) + tv[2] * pv[2]) * inv;
    if (u < 0.0f || u > 1.0f) return 0;
    float qv[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2], tv[0] * e1[1] - tv[1] * e1[0]};
    float v = ((d[0] * qv[0] + d[1] * qv[1]) + d[2] * qv[2]) * inv;
    if (v < 0.0f || u + v > 1.0f) return 0;
    float t = ((e2[0] * qv[0] + e2[1] * qv[1]) + e2[2] * qv[2]) * inv;
    if (!(t > 0.0f)) return 0;
    *t_out = t;
    *entry = det > 0.0f; /* det = -d . (e1 x e2): the ray meets the outside */
    return 1;
}

typedef struct {
    float t;
    int tri;
    int entry;
} hit_t;

static int hit_cmp(const void *a, const void *b) {
    const hit_t *x = (const hit_t *)a, *y = (const hit_t *)b;
    if (x->t < y->t) return -1;
    if (x->t > y->t) return 1;
    return (x->tri > y->tri) - (x->tri < y->tri);
}

static int dup_of(float t_kept, float t) { return t - t_kept <= 1e-4f + 1e-6f * t; }

/* peel_postprocess_kernel.cu:15-158, after its first block (which decodes
 * the dual depth-peeling channel layout into facings): Ts/F hold n entries,
 * padded with INFINITY / 0. */
static void tide(float *Ts, int8_t *F, int n, float sdd) {
    const float cutoff = 0.00001f;
    for (int i = 0; i < n; i++) {
        if (Ts[i] < cutoff || Ts[i] > sdd - 0.001f) {
            Ts[i] = INFINITY;
            F[i] = 0;
        }
    }
    /* selection sort (project order: alpha) */
    for (int s = 0; s < n; s++) {
        int mi = s;
        float mt = Ts[mi];
        for (int i = s + 1; i < n; i++)
            if (Ts[i] < mt) {
                mi = i;
                mt = Ts[i];
            }
        float tt = Ts[s];
        Ts[s] = mt;
        Ts[mi] = tt;
        int8_t tf = F[s];
        F[s] = F[mi];
        F[mi] = tf;
    }
    /* remove duplicates: compare each entry with the last one kept */
    {
        int dst = 0, src = 1;
        while (src < n) {
            int same = (F[src] == F[dst]) &&
                       ((Ts[src] == Ts[dst]) || (F[src] != 0 && dup_of(Ts[dst], Ts[src])));
            if (same) {
                Ts[src] = INFINITY;
                F[src] = 0;
                src++;
            } else {
                dst = src;
                src++;
            }
        }
    }
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) {
            int alt[DRR_MESH_MAX_HITS];
            int a = 0;
            for (int i = 0; i < n; i++) {
                a += F[i];
                alt[i] = a;
            }
            int sea = a > 0 ? a : 0;
            int prev = 0;
            for (int i = 0; i < n; i++) {
                int cur = alt[i];
                if (cur < sea || prev < sea) {
                    Ts[i] = INFINITY;
                    F[i] = 0;
                }
                if (cur > 1 || prev > 1) {
                    Ts[i] = INFINITY;
                    F[i] = 0;
                }
                prev = cur;
            }
        }
        /* fill gaps */
        int dst = 0, src;
        while (dst < n && F[dst] != 0) dst++;
        src = dst + 1;
        while (src < n && dst < n) {
            while (src < n && F[src] == 0) src++;
            if (src < n) {
                Ts[dst] = Ts[src];
                F[dst] = F[src];
                Ts[src] = INFINITY;
                F[src] = 0;
            }
            src++;
            dst++;
        }
    }
}

void ora_mesh_ray_terms(const ora_mesh *m, const float src[3], const float dir[3], float sdd,
                        ora_mesh_ray *out) {
    const int NL = m->n_layers, MH = m->max_hits;
    hit_t *hits = (hit_t *)malloc(sizeof(hit_t) * (size_t)(m->n_triangles > 0 ? m->n_triangles : 1));
    int nh = 0;
    for (int k = 0; k < m->n_triangles; k++) {
        const int32_t *tri = &m->triangles[3 * k];
        float t;
        int entry;
        if (ora_ray_triangle(src, dir, &m->vertices[3 * tri[0]], &m->vertices[3 * tri[1]],
                             &m->vertices[3 * tri[2]], &t, &entry)) {
            hits[nh].t = t;
            hits[nh].tri = k;
            hits[nh].entry = entry;
            nh++;
        }
    }
    qsort(hits, (size_t)nh, sizeof(hit_t), hit_cmp);
    if (nh > DRR_MESH_RAY_HITS) nh = DRR_MESH_RAY_HITS;

    const float sub_far = sdd > 0 ? 2.0f * sdd : INFINITY;  /* kernelTide's argument (projector.py:1227) */
    const float add_far = sdd > 0 ? 4.0f * sdd : INFINITY;  /* zfar of the additive pass (projector.py:871) */
    /* subtractive layers */
    for (int j = 0; j < NL; j++) {
        float *Ts = &out->hit_alpha[j * MH];
        int8_t *F = &out->hit_facing[j * MH];
        int n = 0, any = 0;
        for (int p = 0; p < m->n_prims; p++)
            if (m->prims[p].subtractive && m->prims[p].layer == j) any = 1;
        out->layer_valid[j] = (int8_t)any;
        for (int i = 0; i < nh && n < MH; i++) {
            const drr_mesh_prim_t *pr = &m->prims[m->tri_prim[hits[i].tri]];
            if (!pr->subtractive || pr->layer != j) continue;
            Ts[n] = hits[i].t;
            F[n] = hits[i].entry ? 1 : -1;
            n++;
        }
        for (int i = n; i < MH; i++) {
            Ts[i] = INFINITY;
            F[i] = 0;
        }
        if (any) tide(Ts, F, MH, sub_far);
    }
    /* additive sums per (layer, unique material), then the overlaps of higher
     * subtractive layers (density_between.frag), summed in double */
    for (int L = 0; L < NL; L++) {
        for (int u = 0; u < out->n_unique; u++) {
            const int mat = out->unique_material[u];
            double a = 0.0;
            int fsum = 0;
            float last_t = 0.0f;
            int last_mult = 0;
            for (int i = 0; i < nh; i++) {
                const drr_mesh_prim_t *pr = &m->prims[m->tri_prim[hits[i].tri]];
                if (!pr->additive || pr->layer != L || pr->material != mat) continue;
                if (!(hits[i].t <= add_far)) continue;
                const int mult = hits[i].entry ? -1 : 1;
                if (last_mult == mult && dup_of(last_t, hits[i].t)) continue;
                last_t = hits[i].t;
                last_mult = mult;
                const float rho = pr->density < 0 ? 0.0f : pr->density;
                a += (double)((float)mult * rho) * (double)hits[i].t;
                fsum += mult;
            }
            /* density_between.frag for each higher subtractive layer */
            for (int S = NL - 1; S > L; S--) {
                if (!out->layer_valid[S]) continue;
                const float *Ts = &out->hit_alpha[S * MH];
                for (int p = 0; p < MH / 2; p++) {
                    const float nearD = Ts[2 * p], farD = Ts[2 * p + 1];
                    if (!(farD <= add_far) || isinf(farD)) continue; /* discard (MaxDepth) */
                    last_mult = 0;
                    for (int i = 0; i < nh; i++) {
                        const drr_mesh_prim_t *pr = &m->prims[m->tri_prim[hits[i].tri]];
                        if (!pr->additive || pr->layer != L || pr->material != mat) continue;
                        if (!(hits[i].t <= add_far)) continue;
                        const int mult = hits[i].entry ? -1 : 1;
                        if (last_mult == mult && dup_of(last_t, hits[i].t)) continue;
                        last_t = hits[i].t;
                        last_mult = mult;
                        const double rf = (double)((float)mult * pr->density);
                        const double d = hits[i].t;
                        double r = 0.0;
                        if (d < nearD) r += -(double)nearD * rf;
                        if (d < farD) r += (double)farD * rf;
                        if (d >= nearD && d < farD) r += -d * rf;
                        a += r;
                    }
                }
            }
            out->additive[L * out->n_unique + u] = (float)a;
            out->additive_facing[L * out->n_unique + u] = (float)fsum;
        }
    }
    free(hits);
}
