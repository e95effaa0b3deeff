/*
 * mc_cpu.c -- CPU counterpart of the Woodcock-tracking Monte-Carlo kernel.
 *
 * TEST INFRASTRUCTURE ONLY (the statistical checker of deepdrr_amd/csrc/mc.hip;
 * the product never links it).  deepdrr has no scatter kernel any more
 * (projector.py:527-531); this restates the MC-GPU method the engine
 * implements: point source, spectrum sampling by the cumulative pdf, emission
 * uniform in a cone, Woodcock tracking against a local majorant per
 * 16x16x16-voxel brick (component-wise maxima of density x the three process
 * coefficients; a flight spends its optical depth brick by brick), photoelectric /
 * Compton (Klein-Nishina, Kahn's method) / Rayleigh (Thomson) interactions,
 * scoring on the detector as [primary, single Compton, single Rayleigh,
 * multiple].  Random numbers: RANECU (L'Ecuyer 1988), one stream per history
 * seeded by SplitMix64 of (seed, history) -- the same streams the GPU uses, so
 * histories agree except where the float functions differ by a few ulp (the
 * GPU takes its log, powers, sine and cosine from the hardware log2, exp2,
 * sin and cos) and flip a sampling decision.  Images accumulate in double per
 * thread.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/deepdrr_amd.h"

typedef struct {
    int s1, s2;
} ranecu_t;

static inline uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline ranecu_t seed_history(uint64_t seed, uint64_t h) {
    uint64_t x = seed ^ (h * 0xD1B54A32D192ED03ull);
    ranecu_t r;
    r.s1 = (int)(splitmix64(&x) % 2147483562ull) + 1;
    r.s2 = (int)(splitmix64(&x) % 2147483398ull) + 1;
    return r;
}

static inline float ranecu(ranecu_t *r) {
    int i1 = r->s1 / 53668;
    r->s1 = 40014 * (r->s1 - i1 * 53668) - i1 * 12211;
    if (r->s1 < 0) r->s1 += 2147483563;
    int i2 = r->s2 / 52774;
    r->s2 = 40692 * (r->s2 - i2 * 52774) - i2 * 3791;
    if (r->s2 < 0) r->s2 += 2147483399;
    int z = r->s1 - r->s2;
    if (z < 1) z += 2147483562;
    return fminf((float)z * 4.656612873077393e-10f, 0.99999994f);
}

static void rotate(float *u, float *v, float *w, float ct, float phi) {
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float cp = cosf(phi), sp = sinf(phi);
    if (fabsf(*w) > 0.99999f) {
        *u = st * cp;
        *v = st * sp;
        *w = (*w > 0.0f ? 1.0f : -1.0f) * ct;
    } else {
        float t = sqrtf(1.0f - *w * *w);
        float nu = *u * ct + st * (*u * *w * cp - *v * sp) / t;
        float nv = *v * ct + st * (*v * *w * cp + *u * sp) / t;
        float nw = *w * ct - t * st * cp;
        *u = nu;
        *v = nv;
        *w = nw;
    }
    float n = 1.0f / sqrtf(*u * *u + *v * *v + *w * *w);
    *u *= n;
    *v *= n;
    *w *= n;
}

static float compton(ranecu_t *rng, float *E) {
    float k = *E / 510.99895f;
    float b1 = (1.0f + 2.0f * k) / (9.0f + 2.0f * k);
    float eta = 1.0f, ct;
    for (int it = 0; it < 1000; it++) {
        float r1 = ranecu(rng), r2 = ranecu(rng), r3 = ranecu(rng);
        if (r1 <= b1) {
            eta = 1.0f + 2.0f * k * r2;
            if (r3 <= 4.0f * (1.0f / eta - 1.0f / (eta * eta))) break;
        } else {
            eta = (1.0f + 2.0f * k) / (1.0f + 2.0f * k * r2);
            ct = 1.0f - (eta - 1.0f) / k;
            if (r3 <= 0.5f * (ct * ct + 1.0f / eta)) break;
        }
    }
    ct = 1.0f - (eta - 1.0f) / k;
    *E = *E / eta;
    return fminf(1.0f, fmaxf(-1.0f, ct));
}

static float rayleigh(ranecu_t *rng) {
    float c = 0.0f;
    for (int it = 0; it < 1000; it++) {
        c = 2.0f * ranecu(rng) - 1.0f;
        if (ranecu(rng) <= 0.5f * (1.0f + c * c)) break;
    }
    return c;
}

#define TWO_PI 6.283185307179586f
#define E_MIN 5.0f

static void cross_sections(const drr_mc_params_t *p, float E, float *fpe, float *fc, float *fr) {
    *fpe = powf(E / 60.0f, -p->k_pe);
    *fc = powf(E / 60.0f, -p->k_compton);
    *fr = powf(E / 60.0f, -p->k_rayleigh);
}

#define BRICK 16

typedef struct {
    int nb[3];
    float *maj; /* [nbz][nby][nbx][3]: max rho a_pe, max rho a_c, max rho a_r */
} bricks_t;

static void build_bricks(const float *dens, const uint8_t *lab, const int *n, const drr_mc_params_t *p,
                         bricks_t *b) {
    for (int a = 0; a < 3; a++) b->nb[a] = (n[a] + BRICK - 1) / BRICK;
    size_t nbr = (size_t)b->nb[0] * b->nb[1] * b->nb[2];
    b->maj = (float *)calloc(3 * nbr, sizeof(float));
    for (int k = 0; k < n[2]; k++)
        for (int j = 0; j < n[1]; j++)
            for (int i = 0; i < n[0]; i++) {
                size_t vi = ((size_t)k * n[1] + j) * n[0] + i;
                size_t bi = (((size_t)(k / BRICK) * b->nb[1] + j / BRICK) * b->nb[0] + i / BRICK) * 3;
                float rho = dens[vi];
                int m = lab[vi];
                b->maj[bi] = fmaxf(b->maj[bi], rho * p->a_pe[m]);
                b->maj[bi + 1] = fmaxf(b->maj[bi + 1], rho * p->a_compton[m]);
                b->maj[bi + 2] = fmaxf(b->maj[bi + 2], rho * p->a_rayleigh[m]);
            }
}

/* One flight from x along d with optical depth tau: walks the bricks
 * (Amanatides-Woo) and returns the distance of the tentative collision and its
 * brick majorant in *mub, or -1 if the photon leaves the box (t_exit) first.
 * *bsteps counts the bricks crossed. */
static float flight(const bricks_t *b, const float *A, const float *x, const float *d, float tau, float t_exit,
                    float fpe, float fc, float fr, float *mub_out, long *bsteps) {
    int bi[3], sg[3];
    float tn[3], td[3];
    for (int a = 0; a < 3; a++) {
        float qp = A[4 * a] * x[0] + A[4 * a + 1] * x[1] + A[4 * a + 2] * x[2] + A[4 * a + 3] + 0.5f;
        float dq = A[4 * a] * d[0] + A[4 * a + 1] * d[1] + A[4 * a + 2] * d[2];
        int c = (int)floorf(qp * (1.0f / BRICK));
        bi[a] = c < 0 ? 0 : (c > b->nb[a] - 1 ? b->nb[a] - 1 : c);
        float idq = 1.0f / dq; /* as the kernel: one division per axis */
        if (dq > 0.0f) {
            tn[a] = ((float)(BRICK * (bi[a] + 1)) - qp) * idq;
            td[a] = (float)BRICK * idq;
            sg[a] = 1;
        } else if (dq < 0.0f) {
            tn[a] = ((float)(BRICK * bi[a]) - qp) * idq;
            td[a] = -(float)BRICK * idq;
            sg[a] = -1;
        } else {
            tn[a] = INFINITY;
            td[a] = INFINITY;
            sg[a] = 0;
        }
        tn[a] = fmaxf(tn[a], 0.0f);
    }
    float t = 0.0f;
    for (;;) {
        ++*bsteps;
        const float *mj = &b->maj[(((size_t)bi[2] * b->nb[1] + bi[1]) * b->nb[0] + bi[0]) * 3];
        float mub = (mj[0] * fpe + mj[1] * fc) + mj[2] * fr;
        int a = (tn[0] <= tn[1] && tn[0] <= tn[2]) ? 0 : (tn[1] <= tn[2] ? 1 : 2);
        float tnext = fminf(tn[a], t_exit);
        float seg = tnext - t;
        if (mub > 0.0f) {
            if (tau < mub * seg) {
                *mub_out = mub;
                return t + tau / mub;
            }
            tau -= mub * seg;
        }
        t = tnext;
        if (tnext >= t_exit) return -1.0f;
        bi[a] += sg[a];
        if (bi[a] < 0 || bi[a] >= b->nb[a]) return -1.0f;
        tn[a] += td[a];
    }
}

static float qcoord(const float *A, int a, const float *x) {
    return A[4 * a] * x[0] + A[4 * a + 1] * x[1] + A[4 * a + 2] * x[2] + A[4 * a + 3];
}

static float dcoord(const float *A, int a, const float *d) {
    return A[4 * a] * d[0] + A[4 * a + 1] * d[1] + A[4 * a + 2] * d[2];
}

/* Energy bookkeeping of one history, in keV (ora_mc_tally): where the emitted
 * energy ends.  Indices of tally[]. */
enum { T_EMITTED, T_SCORED, T_PHOTO, T_COMPTON, T_CUTOFF, T_ESCAPED, T_N };

/* One history: returns class (0..3) and energy via *E_out, or -1 if not scored;
 * *steps counts Woodcock flights; tally (may be NULL) accumulates the energy
 * balance of the history. */
static int history(const float *dens, const uint8_t *lab, const int *n, const drr_mc_params_t *p,
                   const bricks_t *bk, uint64_t seed, uint64_t h, float *E_out, int *iu_out, int *iv_out,
                   long *steps, long *bsteps, double *tally) {
    ranecu_t rng = seed_history(seed, h);
    float u = ranecu(&rng);
    int lo = 0, hi = p->n_bins - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (p->cdf[mid] >= u) hi = mid; else lo = mid + 1;
    }
    float E = p->energies_keV[lo];
    if (tally) tally[T_EMITTED] += E;
    float d[3] = {p->cone_axis[0], p->cone_axis[1], p->cone_axis[2]};
    float ct = 1.0f - ranecu(&rng) * (1.0f - p->cone_cos);
    rotate(&d[0], &d[1], &d[2], ct, TWO_PI * ranecu(&rng));
    float x[3] = {p->source_world[0], p->source_world[1], p->source_world[2]};
    int nc = 0, nr = 0;
    const float *A = p->ijk_from_world;
    float tmin = 0.0f, tmax = INFINITY;
    int hit = 1;
    for (int a = 0; a < 3; a++) {
        float q0 = qcoord(A, a, x), dq = dcoord(A, a, d);
        if (dq != 0.0f) {
            float idq = 1.0f / dq;
            float t0 = (-0.5f - q0) * idq, t1 = ((float)n[a] - 0.5f - q0) * idq;
            tmin = fmaxf(tmin, fminf(t0, t1));
            tmax = fminf(tmax, fmaxf(t0, t1));
        } else if (q0 < -0.5f || q0 > (float)n[a] - 0.5f) {
            hit = 0;
        }
    }
    if (hit && tmin < tmax) {
        for (int a = 0; a < 3; a++) x[a] += tmin * d[a];
        float t_exit = tmax - tmin;
        float fpe, fc, fr;
        cross_sections(p, E, &fpe, &fc, &fr);
        for (;;) {
            ++*steps;
            float mub = 0.0f;
            float s = flight(bk, A, x, d, -logf(ranecu(&rng)), t_exit, fpe, fc, fr, &mub, bsteps);
            if (s < 0.0f) {
                for (int a = 0; a < 3; a++) x[a] += t_exit * d[a];
                break;
            }
            for (int a = 0; a < 3; a++) x[a] += s * d[a];
            t_exit -= s;
            int ix[3];
            for (int a = 0; a < 3; a++) {
                int i = (int)floorf(qcoord(A, a, x) + 0.5f);
                ix[a] = i < 0 ? 0 : (i > n[a] - 1 ? n[a] - 1 : i);
            }
            size_t vi = ((size_t)ix[2] * n[1] + ix[1]) * n[0] + ix[0];
            float rho = dens[vi];
            int m = lab[vi];
            float mpe = rho * p->a_pe[m] * fpe;
            float mc = rho * p->a_compton[m] * fc;
            float mr = rho * p->a_rayleigh[m] * fr;
            float mt = mpe + mc + mr;
            if (ranecu(&rng) * mub < mt) {
                float uu = ranecu(&rng) * mt;
                if (uu < mpe) {
                    if (tally) tally[T_PHOTO] += E;
                    return -1;
                }
                if (uu < mpe + mc) {
                    const float E0 = E;
                    ct = compton(&rng, &E);
                    if (tally) tally[T_COMPTON] += (double)E0 - (double)E;
                    nc++;
                } else {
                    ct = rayleigh(&rng);
                    nr++;
                }
                rotate(&d[0], &d[1], &d[2], ct, TWO_PI * ranecu(&rng));
                if (E < E_MIN || nc + nr > p->max_interactions) {
                    if (tally) tally[T_CUTOFF] += E;
                    return -1;
                }
                float tm = INFINITY;
                for (int a = 0; a < 3; a++) {
                    float q0 = qcoord(A, a, x), dq = dcoord(A, a, d);
                    if (dq != 0.0f) {
                        float idq = 1.0f / dq;
                        tm = fminf(tm, fmaxf((-0.5f - q0) * idq, ((float)n[a] - 0.5f - q0) * idq));
                    }
                }
                t_exit = fmaxf(tm, 0.0f);
                cross_sections(p, E, &fpe, &fc, &fr);
            }
        }
    }
    const float *N = p->detector_normal, *O = p->detector_origin, *U = p->detector_u, *V = p->detector_v;
    float dn = d[0] * N[0] + d[1] * N[1] + d[2] * N[2];
    if (dn <= 0.0f) {
        if (tally) tally[T_ESCAPED] += E;
        return -1;
    }
    float t = ((O[0] - x[0]) * N[0] + (O[1] - x[1]) * N[1] + (O[2] - x[2]) * N[2]) / dn;
    if (t < 0.0f) {
        if (tally) tally[T_ESCAPED] += E;
        return -1;
    }
    float r[3];
    for (int a = 0; a < 3; a++) r[a] = x[a] + t * d[a] - O[a];
    float iuu = 1.0f / (U[0] * U[0] + U[1] * U[1] + U[2] * U[2]);
    float ivv = 1.0f / (V[0] * V[0] + V[1] * V[1] + V[2] * V[2]);
    float fu = (r[0] * U[0] + r[1] * U[1] + r[2] * U[2]) * iuu;
    float fv = (r[0] * V[0] + r[1] * V[1] + r[2] * V[2]) * ivv;
    if (!(fu >= 0.0f && fv >= 0.0f && fu < (float)p->det_width && fv < (float)p->det_height)) {
        if (tally) tally[T_ESCAPED] += E;
        return -1;
    }
    if (tally) tally[T_SCORED] += E;
    int iu = (int)fu, iv = (int)fv;
    if (iu > p->det_width - 1) iu = p->det_width - 1;
    if (iv > p->det_height - 1) iv = p->det_height - 1;
    *E_out = E;
    *iu_out = iu;
    *iv_out = iv;
    return (nc + nr == 0) ? 0 : (nc == 1 && nr == 0) ? 1 : (nr == 1 && nc == 0) ? 2 : 3;
}

/* dens/lab [nz][ny][nx]; img/img2 double [4][H][W] (accumulated);
 * returns the number of Woodcock flights; *brick_steps (may be NULL)
 * receives the number of brick majorants they crossed. */
long ora_mc(const float *dens, const uint8_t *lab, int nx, int ny, int nz, const drr_mc_params_t *p,
            int64_t first, int64_t count, uint64_t seed, double *img, double *img2, int n_threads,
            long *brick_steps) {
    const int n[3] = {nx, ny, nz};
    const long npix = (long)p->det_width * p->det_height;
    bricks_t bk;
    build_bricks(dens, lab, n, p, &bk);
    long total_steps = 0, total_bsteps = 0;
#pragma omp parallel num_threads(n_threads) reduction(+ : total_steps, total_bsteps)
    {
        double *li = (double *)calloc(4 * npix, sizeof(double));
        double *l2 = (double *)calloc(4 * npix, sizeof(double));
        long steps = 0, bsteps = 0;
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < count; i++) {
            float E;
            int iu, iv;
            int cls = history(dens, lab, n, p, &bk, seed, (uint64_t)(first + i), &E, &iu, &iv, &steps, &bsteps,
                              NULL);
            if (cls >= 0) {
                long o = cls * npix + (long)iv * p->det_width + iu;
                li[o] += E;
                l2[o] += (double)E * E;
            }
        }
#pragma omp critical
        {
            for (long k = 0; k < 4 * npix; k++) {
                img[k] += li[k];
                img2[k] += l2[k];
            }
        }
        free(li);
        free(l2);
        total_steps += steps;
        total_bsteps += bsteps;
    }
    free(bk.maj);
    if (brick_steps) *brick_steps = total_bsteps;
    return total_steps;
}

/* Samplers exposed for distribution tests: n draws of Compton (cos, E') at
 * energy E, and of Rayleigh cos, from one RANECU stream. */
void ora_sample_compton(float E, int n, uint64_t seed, float *cos_out, float *e_out) {
    ranecu_t r = seed_history(seed, 0);
    for (int i = 0; i < n; i++) {
        float e = E;
        cos_out[i] = compton(&r, &e);
        e_out[i] = e;
    }
}

void ora_sample_rayleigh(int n, uint64_t seed, float *cos_out) {
    ranecu_t r = seed_history(seed, 0);
    for (int i = 0; i < n; i++) cos_out[i] = rayleigh(&r);
}

/* Energy balance of histories [first, first + count): tally[6] receives, in
 * keV, the emitted energy and where it ended -- scored on the detector,
 * absorbed by photoelectric events, deposited by Compton recoils, dropped by
 * the energy / interaction cutoff, escaped past the detector.  counts[4]
 * receives the scored histories per class.  Single-threaded. */
long ora_mc_tally(const float *dens, const uint8_t *lab, int nx, int ny, int nz, const drr_mc_params_t *p,
                  int64_t first, int64_t count, uint64_t seed, double *tally, long *counts) {
    const int n[3] = {nx, ny, nz};
    long steps = 0, bsteps = 0;
    bricks_t bk;
    build_bricks(dens, lab, n, p, &bk);
    for (int k = 0; k < T_N; k++) tally[k] = 0.0;
    for (int k = 0; k < 4; k++) counts[k] = 0;
    for (int64_t i = 0; i < count; i++) {
        float E;
        int iu, iv;
        int cls = history(dens, lab, n, p, &bk, seed, (uint64_t)(first + i), &E, &iu, &iv, &steps, &bsteps, tally);
        if (cls >= 0) counts[cls]++;
    }
    free(bk.maj);
    return steps;
}

/* The direction rotation of an interaction: (u, v, w) turned by polar cosine
 * ct and azimuth phi (out[3]). */
void ora_rotate(float u, float v, float w, float ct, float phi, float *out) {
    rotate(&u, &v, &w, ct, phi);
    out[0] = u;
    out[1] = v;
    out[2] = w;
}

/* The first n draws of history h's RANECU stream: the two generator states
 * after each draw and the uniform it returned. */
void ora_ranecu(uint64_t seed, uint64_t h, int n, int *s1, int *s2, float *u) {
    ranecu_t r = seed_history(seed, h);
    for (int i = 0; i < n; i++) {
        u[i] = ranecu(&r);
        s1[i] = r.s1;
        s2[i] = r.s2;
    }
}

/* The seeds of history h's stream (before any draw). */
void ora_ranecu_seed(uint64_t seed, uint64_t h, int *s1, int *s2) {
    ranecu_t r = seed_history(seed, h);
    *s1 = r.s1;
    *s2 = r.s2;
}
