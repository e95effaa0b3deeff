/*
 * transcription.c -- CPU transcription of deepdrr's projectKernel.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle for the HIP
 * kernels in deepdrr_amd/csrc; the product never links or calls it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * It follows deepdrr/projector/project_kernel.cu:135-650 statement by
 * statement in FP32 with the same operation order, built with
 * -ffp-contract=off (no implicit FMA contraction) and without fast-math.
 * Three multiply-adds are fused explicitly with fmaf, where nvcc's default
 * --fmad=true fuses the reference's expressions: the sample position
 * s + alpha * r - 1 (402-404), the accumulation ad[m] += (w * rho) * seg[m]
 * (543-547) and the running sum of the exact-weight trilinear filter below.
 *   ray setup            project_kernel.cu:195-234
 *   slab clip            project_kernel.cu:242-323
 *   step count           project_kernel.cu:334
 *   air prefix           project_kernel.cu:359-361 (double promotion of 0.1129)
 *   march                project_kernel.cu:398-553
 *     label cache shared across volumes     394-431
 *     exact FP32 trilinear weights          434-448
 *     soft material fractions               451-455
 *     priority selection                    458-496
 *     accumulation with 1/n and end halves  522-549
 *   air suffix, step scale, /10             557-584
 *   spectral Beer-Lambert sum               637-646
 *   meshes (ora_project_mesh): hit lists    363-375
 *     inside-mesh test per step             498-521
 *     additive mesh densities               569-579
 *     (the hit lists themselves: oracle/mesh_hits.c)
 *
 * Texture semantics, restated exactly (the reference samples CUDA textures):
 *   labels  (point, clamp): corner d of the cell is L[clamp(floor(p + d))]
 *   density (linear, clamp, fetched at p + 0.5): corners clamp(floor(p)) and
 *           clamp(floor(p) + 1) weighted by the same exact FP32 weights as the
 *           labels, summed in (dz, dy, dx) order starting from the first
 *           product, each further term fused (fmaf).  (NVIDIA's texture unit uses
 *           8-bit fixed-point weights; exact weights are this engine's
 *           definition.)
 *   The label cache starts empty (the reference's first fetch is skipped if
 *   the first base is (-1,-1,-1), leaving it uninitialised), and disabled
 *   volumes neither fetch nor touch the cache (the reference reads an
 *   uninitialised ray direction for them).
 * Output index: row-major [v][u] (the reference's [u][v] after the host swap).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/deepdrr_amd.h"
#include "mesh_hits.h"

#define AIR_DENSITY 0.1129 /* double literal, as project_kernel.cu:6 */

typedef struct {
    const float *density; /* [nz][ny][nx] */
    const uint8_t *labels;
    int n[3];
} ora_volume;

static inline int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

static inline int clamp_index(float x, int n) {
    /* floor already applied; clamp in float first so huge values stay defined */
    if (!(x >= 0.0f)) return 0;             /* also NaN */
    if (x > (float)(n - 1)) return n - 1;
    return (int)x;
}

static inline uint8_t label_at(const ora_volume *v, float x, float y, float z) {
    int i = clamp_index(floorf(x), v->n[0]);
    int j = clamp_index(floorf(y), v->n[1]);
    int k = clamp_index(floorf(z), v->n[2]);
    return v->labels[((size_t)k * v->n[1] + j) * v->n[0] + i];
}

static inline float density_at(const ora_volume *v, int i, int j, int k) {
    i = clampi(i, 0, v->n[0] - 1);
    j = clampi(j, 0, v->n[1] - 1);
    k = clampi(k, 0, v->n[2] - 1);
    return v->density[((size_t)k * v->n[1] + j) * v->n[0] + i];
}

static inline int floor_to_int(float b) {
    /* b is floorf(p); clamp to a safe int range before converting */
    if (!(b > -2.0e9f)) return -2000000000;
    if (b > 2.0e9f) return 2000000000;
    return (int)b;
}

/* One ray.  ad[] receives the final area densities (g/cm^2). */
typedef struct {
    const ora_mesh *mesh;
    float src[3]; /* source, world */
    float sdd;
    int n_unique;
    int unique_material[DRR_MAX_MATERIALS];
} ora_mesh_scene;

static void ora_ray(const ora_volume *vols, const drr_view_t *view, const drr_scene_t *sc,
                    int udx, int vdx, int n_bins, const float *energies, const float *pdf,
                    const float *mu, float *out_intensity, float *out_photon_prob, float *ad,
                    const ora_mesh_scene *ms) {
    const int V = view->n_volumes;
    const int M = sc->n_materials;
    const float step = sc->step;
    const float *W = view->world_from_index;

    float u = (float)((double)(float)udx + 0.5);
    float v = (float)((double)(float)vdx + 0.5);
    float rx = u * W[0] + v * W[1] + W[2];
    float ry = u * W[3] + v * W[4] + W[5];
    float rz = u * W[6] + v * W[7] + W[8];
    float ray_length = sqrtf(rx * rx + ry * ry + rz * rz);
    float inv_ray_norm = 1.0f / ray_length;
    rx *= inv_ray_norm;
    ry *= inv_ray_norm;
    rz *= inv_ray_norm;

    float minAlpha = ray_length;
    float maxAlpha = 0;
    for (int m = 0; m < M; m++) ad[m] = 0.0f;

    /* the meshes' per-ray terms (what OpenGL and the peel kernels hand to the
     * reference kernel), and its per-layer hit cursors (363-375) */
    ora_mesh_ray mr;
    int NL = 0, MH = 0;
    int mesh_hit_depth[DRR_MESH_MAX_LAYERS], mesh_hit_index[DRR_MESH_MAX_LAYERS];
    if (ms) {
        const float dir[3] = {rx, ry, rz};
        mr.n_unique = ms->n_unique;
        for (int i = 0; i < ms->n_unique; i++) mr.unique_material[i] = ms->unique_material[i];
        ora_mesh_ray_terms(ms->mesh, ms->src, dir, ms->sdd, &mr);
        NL = ms->mesh->n_layers;
        MH = ms->mesh->max_hits;
        for (int j = 0; j < NL; j++) {
            mesh_hit_depth[j] = 0;
            mesh_hit_index[j] = 0;
        }
    }

    float minA_v[DRR_MAX_VOLUMES], maxA_v[DRR_MAX_VOLUMES];
    int do_trace[DRR_MAX_VOLUMES];
    float rxi[DRR_MAX_VOLUMES], ryi[DRR_MAX_VOLUMES], rzi[DRR_MAX_VOLUMES];
    for (int i = 0; i < V; i++) {
        const drr_view_volume_t *vv = &view->vol[i];
        rxi[i] = ryi[i] = rzi[i] = 0.0f;
        if (vv->enabled == 0) {
            do_trace[i] = 0;
            continue;
        }
        const float *A = vv->ijk_from_world;
        rxi[i] = A[0] * rx + A[1] * ry + A[2] * rz + A[3] * 0;
        ryi[i] = A[4] * rx + A[5] * ry + A[6] * rz + A[7] * 0;
        rzi[i] = A[8] * rx + A[9] * ry + A[10] * rz + A[11] * 0;
        do_trace[i] = 1;
        minA_v[i] = 0;
        maxA_v[i] = view->max_ray_length > 0 ? view->max_ray_length : INFINITY;
        const float r3[3] = {rxi[i], ryi[i], rzi[i]};
        int skip = 0;
        for (int a = 0; a < 3; a++) {
            if (0.0f != r3[a]) {
                float reci = 1.0f / r3[a];
                float alpha0 = (vv->bounds_lo[a] - vv->source_ijk[a]) * reci;
                float alpha1 = (vv->bounds_hi[a] - vv->source_ijk[a]) * reci;
                minA_v[i] = fmaxf(minA_v[i], fminf(alpha0, alpha1));
                maxA_v[i] = fminf(maxA_v[i], fmaxf(alpha0, alpha1));
            } else if (vv->bounds_lo[a] > vv->source_ijk[a] || vv->source_ijk[a] > vv->bounds_hi[a]) {
                do_trace[i] = 0;
                skip = 1;
                break;
            }
        }
        if (skip) continue;
        minAlpha = fminf(minAlpha, minA_v[i]);
        maxAlpha = fmaxf(maxAlpha, maxA_v[i]);
    }

    int num_steps = (int)ceilf((maxAlpha - minAlpha) / step);

    if (sc->attenuate_outside) {
        ad[sc->air_index] = (float)((double)ad[sc->air_index] + (double)(minAlpha / step) * AIR_DENSITY);
    }

    float px[DRR_MAX_VOLUMES], py[DRR_MAX_VOLUMES], pz[DRR_MAX_VOLUMES];
    float seg[DRR_MAX_VOLUMES][DRR_MAX_MATERIALS];
    int voxels[2][2][2];
    float weights[2][2][2];
    float prev[3] = {0, 0, 0};
    int have_prev = 0;
    float alpha = minAlpha;

    for (int t = 0; t < num_steps; t++) {
        for (int vi = 0; vi < V; vi++) {
            if (view->vol[vi].enabled == 0) continue;
            const float *s = view->vol[vi].source_ijk;
            px[vi] = fmaf(alpha, rxi[vi], s[0]) - 1.0f;
            py[vi] = fmaf(alpha, ryi[vi], s[1]) - 1.0f;
            pz[vi] = fmaf(alpha, rzi[vi], s[2]) - 1.0f;
            for (int m = 0; m < M; m++) seg[vi][m] = 0.0f;
            float bx = floorf(px[vi]), by = floorf(py[vi]), bz = floorf(pz[vi]);
            if (!have_prev || bx != prev[0] || by != prev[1] || bz != prev[2]) {
                for (int dz = 0; dz <= 1; ++dz)
                    for (int dy = 0; dy <= 1; ++dy)
                        for (int dx = 0; dx <= 1; ++dx)
                            voxels[dx][dy][dz] = label_at(&vols[vi], px[vi] + dx, py[vi] + dy, pz[vi] + dz);
                prev[0] = bx;
                prev[1] = by;
                prev[2] = bz;
                have_prev = 1;
            }
            float fx = px[vi] - bx, fy = py[vi] - by, fz = pz[vi] - bz;
            float fxi = 1.0f - fx, fyi = 1.0f - fy, fzi = 1.0f - fz;
            weights[0][0][0] = fxi * fyi * fzi;
            weights[1][0][0] = fx * fyi * fzi;
            weights[0][1][0] = fxi * fy * fzi;
            weights[1][1][0] = fx * fy * fzi;
            weights[0][0][1] = fxi * fyi * fz;
            weights[1][0][1] = fx * fyi * fz;
            weights[0][1][1] = fxi * fy * fz;
            weights[1][1][1] = fx * fy * fz;
            for (int dz = 0; dz <= 1; ++dz)
                for (int dy = 0; dy <= 1; ++dy)
                    for (int dx = 0; dx <= 1; ++dx)
                        seg[vi][voxels[dx][dy][dz]] += weights[dx][dy][dz];
        }

        int curr_priority = V;
        int n_at = 0;
        for (int i = 0; i < V; i++) {
            if (0 == do_trace[i]) continue;
            if (view->vol[i].enabled == 0) continue;
            if ((alpha < minA_v[i]) || (alpha > maxA_v[i])) continue;
            float any_seg = 0.0f;
            for (int m = 0; m < M; m++) {
                any_seg += seg[i][m];
                if (any_seg > 0.0f) break;
            }
            if (0.0f == any_seg) continue;
            int pr = view->vol[i].priority;
            if (pr < curr_priority) {
                curr_priority = pr;
                n_at = 1;
            } else if (pr == curr_priority) {
                n_at += 1;
            }
        }

        int inside_mesh = 0;
        for (int j = 0; j < NL; j++) {
            if (mr.layer_valid[j] == 0) continue;
            while (mesh_hit_index[j] < MH && mr.hit_facing[j * MH + mesh_hit_index[j]] != 0 &&
                   mr.hit_alpha[j * MH + mesh_hit_index[j]] < alpha) {
                mesh_hit_depth[j] += mr.hit_facing[j * MH + mesh_hit_index[j]];
                mesh_hit_index[j] += 1;
            }
        }
        for (int j = 0; j < NL; j++)
            if (mesh_hit_depth[j] > 0) inside_mesh = 1;

        if (inside_mesh) {
            /* the volumes (and the outside air) are hidden inside a mesh */
        } else if (0 == n_at) {
            if (sc->attenuate_outside)
                ad[sc->air_index] = (float)((double)ad[sc->air_index] + AIR_DENSITY);
        } else {
            float weight = 1.0f / ((float)n_at);
            weight *= (0 == t || num_steps - 1 == t) ? 0.5f : 1.0f;
            for (int vi = 0; vi < V; vi++) {
                if (do_trace[vi] && (view->vol[vi].priority == curr_priority) && (view->vol[vi].enabled == 1)) {
                    /* linear filter at p + 0.5: corners floor(p) and floor(p)+1 */
                    float bx = floorf(px[vi]), by = floorf(py[vi]), bz = floorf(pz[vi]);
                    int ix = floor_to_int(bx), iy = floor_to_int(by), iz = floor_to_int(bz);
                    float fx = px[vi] - bx, fy = py[vi] - by, fz = pz[vi] - bz;
                    float fxi = 1.0f - fx, fyi = 1.0f - fy, fzi = 1.0f - fz;
                    float w[8];
                    w[0] = fxi * fyi * fzi;
                    w[1] = fx * fyi * fzi;
                    w[2] = fxi * fy * fzi;
                    w[3] = fx * fy * fzi;
                    w[4] = fxi * fyi * fz;
                    w[5] = fx * fyi * fz;
                    w[6] = fxi * fy * fz;
                    w[7] = fx * fy * fz;
                    float rho = w[0] * density_at(&vols[vi], ix, iy, iz);
                    for (int d = 1; d < 8; d++) {
                        int dx = d & 1, dy = (d >> 1) & 1, dz = d >> 2;
                        rho = fmaf(w[d], density_at(&vols[vi], ix + dx, iy + dy, iz + dz), rho);
                    }
                    for (int m = 0; m < M; m++) ad[m] = fmaf(weight * rho, seg[vi][m], ad[m]);
                }
            }
        }
        alpha += step;
    }

    if (sc->attenuate_outside) {
        ad[sc->air_index] = (float)((double)ad[sc->air_index] +
                                    (double)((ray_length - maxAlpha) / step) * AIR_DENSITY);
    }
    for (int m = 0; m < M; m++) ad[m] *= step;
    if (ms) {
        for (int i = 0; i < mr.n_unique; i++)
            for (int j = 0; j < NL; j++)
                if (fabsf(mr.additive_facing[j * mr.n_unique + i]) < 0.00001f)
                    ad[mr.unique_material[i]] += fmaxf(mr.additive[j * mr.n_unique + i], 0.0f);
    }
    for (int m = 0; m < M; m++) ad[m] /= 10.0f;

    float intensity = 0.0f, photon_prob = 0.0f;
    for (int b = 0; b < n_bins; b++) {
        float e = 0.0f;
        for (int m = 0; m < M; m++) e += ad[m] * mu[b * M + m];
        float p = expf(-1.f * e) * pdf[b];
        photon_prob += p;
        intensity += energies[b] * p;
    }
    *out_intensity = intensity;
    *out_photon_prob = photon_prob;
}

/*
 * Project detector rows of one view.  density[v], labels[v]: [nz][ny][nx] with
 * dims[3v..3v+2] = (nx, ny, nz).  rows: NULL for every row (outputs [H][W]),
 * else n_rows row indices (outputs [n_rows][W], row r of the output is
 * detector row rows[r]); area_density (may be NULL) is [M][rows][W].  Each
 * ray is computed exactly as in the full view.  Rays are split over
 * `n_threads` OpenMP threads (dynamic schedule).
 */
static int ora_project_impl(const float *const *density, const uint8_t *const *labels, const int *dims,
                            const drr_view_t *view, const drr_scene_t *scene, int n_bins,
                            const float *energies, const float *pdf, const float *mu, const int *rows,
                            int n_rows, float *intensity, float *photon_prob, float *area_density,
                            int n_threads, const ora_mesh_scene *ms) {
    if (view->n_volumes < (ms ? 0 : 1) || view->n_volumes > DRR_MAX_VOLUMES) return -1;
    if (scene->n_materials < 1 || scene->n_materials > DRR_MAX_MATERIALS) return -1;
    ora_volume vols[DRR_MAX_VOLUMES];
    for (int v = 0; v < view->n_volumes; v++) {
        vols[v].density = density[v];
        vols[v].labels = labels[v];
        vols[v].n[0] = dims[3 * v];
        vols[v].n[1] = dims[3 * v + 1];
        vols[v].n[2] = dims[3 * v + 2];
    }
    const int Wd = view->width, Hd = view->height, M = scene->n_materials;
    const int Ho = rows ? n_rows : Hd;
    if (rows)
        for (int r = 0; r < n_rows; r++)
            if (rows[r] < 0 || rows[r] >= Hd) return -1;
    const long npix = (long)Wd * Ho;
#pragma omp parallel for schedule(dynamic, 16) num_threads(n_threads)
    for (long idx = 0; idx < npix; idx++) {
        int r = (int)(idx / Wd), udx = (int)(idx % Wd);
        int vdx = rows ? rows[r] : r;
        float ad[DRR_MAX_MATERIALS];
        ora_ray(vols, view, scene, udx, vdx, n_bins, energies, pdf, mu, &intensity[idx],
                &photon_prob[idx], ad, ms);
        if (area_density)
            for (int m = 0; m < M; m++) area_density[(long)m * npix + idx] = ad[m];
    }
    return 0;
}

int ora_project(const float *const *density, const uint8_t *const *labels, const int *dims,
                const drr_view_t *view, const drr_scene_t *scene, int n_bins,
                const float *energies, const float *pdf, const float *mu, float *intensity,
                float *photon_prob, float *area_density, int n_threads) {
    return ora_project_impl(density, labels, dims, view, scene, n_bins, energies, pdf, mu, NULL, 0,
                            intensity, photon_prob, area_density, n_threads, NULL);
}

int ora_project_rows(const float *const *density, const uint8_t *const *labels, const int *dims,
                     const drr_view_t *view, const drr_scene_t *scene, int n_bins,
                     const float *energies, const float *pdf, const float *mu, const int *rows,
                     int n_rows, float *intensity, float *photon_prob, float *area_density,
                     int n_threads) {
    if (!rows || n_rows < 1) return -1;
    return ora_project_impl(density, labels, dims, view, scene, n_bins, energies, pdf, mu, rows, n_rows,
                            intensity, photon_prob, area_density, n_threads, NULL);
}

/*
 * Project one view with meshes (rows: NULL for all, as ora_project_rows).
 * view->n_volumes may be 0.  The mesh's unique materials are those of its
 * additive primitives, in ascending global index.
 */
int ora_project_mesh(const float *const *density, const uint8_t *const *labels, const int *dims,
                     const drr_view_t *view, const drr_scene_t *scene, int n_bins, const float *energies,
                     const float *pdf, const float *mu, const float *vertices, const int32_t *triangles,
                     const int32_t *tri_prim, int n_triangles, const drr_mesh_prim_t *prims, int n_prims,
                     int n_layers, int max_hits, const float *source_world, float sdd, const int *rows,
                     int n_rows, float *intensity, float *photon_prob, float *area_density, int n_threads) {
    if (n_layers < 1 || n_layers > DRR_MESH_MAX_LAYERS || max_hits < 4 || max_hits > DRR_MESH_MAX_HITS ||
        max_hits % 4 != 0)
        return -1;
    ora_mesh mesh = {vertices, triangles, tri_prim, n_triangles, prims, n_prims, n_layers, max_hits};
    ora_mesh_scene ms;
    ms.mesh = &mesh;
    for (int a = 0; a < 3; a++) ms.src[a] = source_world[a];
    ms.sdd = sdd;
    ms.n_unique = 0;
    for (int m = 0; m < scene->n_materials; m++) {
        int used = 0;
        for (int p = 0; p < n_prims; p++)
            if (prims[p].additive && prims[p].material == m) used = 1;
        if (used) ms.unique_material[ms.n_unique++] = m;
    }
    return ora_project_impl(density, labels, dims, view, scene, n_bins, energies, pdf, mu, rows, n_rows,
                            intensity, photon_prob, area_density, n_threads, &ms);
}

/* Per-ray mesh terms of one view for inspection: hidden alpha ranges
 * (a, b] merged over the layers ([H][W][n_layers * max_hits / 2][2], count
 * [H][W]) and the clamped additive terms ([n_layers * n_unique][H][W], 0 where
 * the facing sum is not 0).  Returns n_unique. */
int ora_mesh_rays(const drr_view_t *view, const float *vertices, const int32_t *triangles,
                  const int32_t *tri_prim, int n_triangles, const drr_mesh_prim_t *prims, int n_prims,
                  int n_layers, int max_hits, int n_materials, const float *source_world, float sdd,
                  float *intervals, int32_t *counts, float *additive) {
    ora_mesh mesh = {vertices, triangles, tri_prim, n_triangles, prims, n_prims, n_layers, max_hits};
    int n_unique = 0, uniq[DRR_MAX_MATERIALS];
    for (int m = 0; m < n_materials; m++) {
        int used = 0;
        for (int p = 0; p < n_prims; p++)
            if (prims[p].additive && prims[p].material == m) used = 1;
        if (used) uniq[n_unique++] = m;
    }
    const int W = view->width, H = view->height, K = n_layers * max_hits / 2;
    const long npix = (long)W * H;
    for (long idx = 0; idx < npix; idx++) {
        int vdx = (int)(idx / W), udx = (int)(idx % W);
        const float *Wm = view->world_from_index;
        float u = (float)((double)(float)udx + 0.5);
        float v = (float)((double)(float)vdx + 0.5);
        float rx = u * Wm[0] + v * Wm[1] + Wm[2];
        float ry = u * Wm[3] + v * Wm[4] + Wm[5];
        float rz = u * Wm[6] + v * Wm[7] + Wm[8];
        float L = sqrtf(rx * rx + ry * ry + rz * rz);
        float inv = 1.0f / L;
        const float dir[3] = {rx * inv, ry * inv, rz * inv};
        ora_mesh_ray mr;
        mr.n_unique = n_unique;
        for (int i = 0; i < n_unique; i++) mr.unique_material[i] = uniq[i];
        ora_mesh_ray_terms(&mesh, source_world, dir, sdd, &mr);
        /* inside(alpha) = some layer's depth (sum of facings of hits < alpha) > 0 */
        float ts[DRR_MESH_MAX_LAYERS * DRR_MESH_MAX_HITS];
        int nt = 0;
        for (int j = 0; j < n_layers; j++)
            if (mr.layer_valid[j])
                for (int k = 0; k < max_hits && mr.hit_facing[j * max_hits + k] != 0; k++)
                    ts[nt++] = mr.hit_alpha[j * max_hits + k];
        /* sort the distinct hit alphas; the state is constant between them */
        for (int a = 1; a < nt; a++)
            for (int b = a; b > 0 && ts[b - 1] > ts[b]; b--) {
                float x = ts[b];
                ts[b] = ts[b - 1];
                ts[b - 1] = x;
            }
        int n_iv = 0, open = 0;
        float start = 0.0f;
        for (int k = 0; k < nt; k++) {
            if (k > 0 && ts[k] == ts[k - 1]) continue;
            /* state for alpha just above ts[k] */
            int in = 0;
            for (int j = 0; j < n_layers && !in; j++) {
                if (!mr.layer_valid[j]) continue;
                int d = 0;
                for (int q = 0; q < max_hits && mr.hit_facing[j * max_hits + q] != 0; q++)
                    if (mr.hit_alpha[j * max_hits + q] <= ts[k]) d += mr.hit_facing[j * max_hits + q];
                if (d > 0) in = 1;
            }
            if (in && !open) {
                start = ts[k];
                open = 1;
            } else if (!in && open) {
                intervals[(idx * K + n_iv) * 2] = start;
                intervals[(idx * K + n_iv) * 2 + 1] = ts[k];
                n_iv++;
                open = 0;
            }
        }
        if (open) {
            intervals[(idx * K + n_iv) * 2] = start;
            intervals[(idx * K + n_iv) * 2 + 1] = INFINITY;
            n_iv++;
        }
        counts[idx] = n_iv;
        for (int j = 0; j < n_layers; j++)
            for (int i = 0; i < n_unique; i++)
                additive[(long)(j * n_unique + i) * npix + idx] =
                    fabsf(mr.additive_facing[j * n_unique + i]) < 0.00001f ? fmaxf(mr.additive[j * n_unique + i], 0.0f)
                                                                           : 0.0f;
    }
    return n_unique;
}

/* Solid angle of each pixel seen from the source (project_kernel.cu:14-133),
 * in FP32 with the reference's operation order; output [H][W]. */
void ora_solid_angle(const float *W, int width, int height, float *out) {
    for (int vdx = 0; vdx < height; vdx++)
        for (int udx = 0; udx < width; udx++) {
            float cx[4], cy[4], cz[4], cm[4];
            const float cu_off[4] = {0.f, 1.f, 1.f, 0.f};
            const float cv_off[4] = {0.f, 0.f, 1.f, 1.f};
            for (int c = 0; c < 4; c++) {
                float cu = udx + cu_off[c];
                float cv = vdx + cv_off[c];
                cx[c] = cu * W[0] + cv * W[1] + W[2];
                cy[c] = cu * W[3] + cv * W[4] + W[5];
                cz[c] = cu * W[6] + cv * W[7] + W[8];
                cm[c] = sqrtf((cx[c] * cx[c]) + (cy[c] * cy[c]) + (cz[c] * cz[c]));
            }
            float x02 = (cy[0] * cz[2]) - (cz[0] * cy[2]);
            float y02 = (cz[0] * cx[2]) - (cx[0] * cz[2]);
            float z02 = (cx[0] * cy[2]) - (cy[0] * cx[2]);
            float d01 = (cx[0] * cx[1]) + (cy[0] * cy[1]) + (cz[0] * cz[1]);
            float d02 = (cx[0] * cx[2]) + (cy[0] * cy[2]) + (cz[0] * cz[2]);
            float d03 = (cx[0] * cx[3]) + (cy[0] * cy[3]) + (cz[0] * cz[3]);
            float d12 = (cx[1] * cx[2]) + (cy[1] * cy[2]) + (cz[1] * cz[2]);
            float d23 = (cx[2] * cx[3]) + (cy[2] * cy[3]) + (cz[2] * cz[3]);
            float n012 = fabsf((cx[1] * x02) + (cy[1] * y02) + (cz[1] * z02));
            float n023 = fabsf((cx[3] * x02) + (cy[3] * y02) + (cz[3] * z02));
            float den012 = (cm[0] * cm[1] * cm[2]) + (d01 * cm[2]) + (d02 * cm[1]) + (d12 * cm[0]);
            float den023 = (cm[0] * cm[2] * cm[3]) + (d02 * cm[3]) + (d03 * cm[2]) + (d23 * cm[0]);
            float s012 = 2.f * atan2f(n012, den012);
            if (s012 < 0.0f) s012 += 3.141592654f;
            float s023 = 2.f * atan2f(n023, den023);
            if (s023 < 0.0f) s023 += 3.141592654f;
            out[(long)vdx * width + udx] = s012 + s023;
        }
}
