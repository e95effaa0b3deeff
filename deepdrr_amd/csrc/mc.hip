// mc.hip -- Monte-Carlo X-ray transport with Woodcock (delta) tracking.
//
// deepdrr removed its scatter simulation (projector.py:527-531 raises for
// scatter_num > 0) and keeps only MC-GPU's data files and a RITA sampler
// (mcgpu_*.py, rita.py).  This kernel implements the MC-GPU method (Badal &
// Badano 2009): photons from a point source, energies from the spectrum,
// Woodcock tracking through the voxel phantom against a majorant attenuation
// coefficient, photoelectric absorption, Compton scattering (Klein-Nishina,
// Kahn's rejection sampling) and Rayleigh scattering (Thomson angular
// distribution), and scoring of the energy reaching each detector pixel split
// into primary / single Compton / single Rayleigh / multiple scatter.
// Cross sections are power laws mu_x = rho * a_x[m] * (E / 60 keV)^(-k_x).
// Random numbers: RANECU (L'Ecuyer 1988, as in PENELOPE and MC-GPU), one
// stream per history seeded from (seed, history index) by SplitMix64.
//
// Local majorants: one global majorant (the densest voxel's attenuation, here
// a titanium rod) made a photon in soft tissue take ~155 flights, most of them
// virtual.  The volume is cut into 16x16x16-voxel bricks, and each brick stores
// the component-wise maxima of density x (photoelectric, Compton, Rayleigh)
// coefficient over its voxels (k_mc_bricks, per call); its majorant at energy
// E is maj_pe f_pe(E) + maj_c f_c(E) + maj_r f_r(E), which bounds every
// voxel's attenuation in FP32 too (products and sums round monotonically).  A
// flight draws its optical depth once and walks the bricks along the
// direction (Amanatides-Woo), spending depth at each brick's own majorant;
// where it runs out is the tentative collision, accepted as real with
// probability mu(voxel) / mu(brick).  This is delta tracking with a piecewise
// constant majorant, unbiased like the global one (Badal & Badano 2009 use a
// global majorant; the local form is standard in Woodcock trackers).
//
// GPU mapping: every lane owns one history at a time and starts the next one
// as soon as its photon is absorbed or scored (no waiting for the slowest
// history of the wave); each loop iteration is one flight.
#include <hip/hip_runtime.h>

#include "drr_internal.h"

namespace {

struct Ranecu {
    int s1, s2;
};

__device__ __forceinline__ uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ Ranecu seed_history(uint64_t seed, uint64_t h) {
    uint64_t x = seed ^ (h * 0xD1B54A32D192ED03ull);
    Ranecu r;
    r.s1 = (int)(splitmix64(x) % 2147483562ull) + 1;
    r.s2 = (int)(splitmix64(x) % 2147483398ull) + 1;
    return r;
}

// RANECU: two multiplicative congruential generators (Schrage's method), combined.
__device__ __forceinline__ float ranecu(Ranecu &r) {
    int i1 = r.s1 / 53668;
    r.s1 = 40014 * (r.s1 - i1 * 53668) - i1 * 12211;
    if (r.s1 < 0) r.s1 += 2147483563;
    int i2 = r.s2 / 52774;
    r.s2 = 40692 * (r.s2 - i2 * 52774) - i2 * 3791;
    if (r.s2 < 0) r.s2 += 2147483399;
    int z = r.s1 - r.s2;
    if (z < 1) z += 2147483562;
    return fminf((float)z * 4.656612873077393e-10f, 0.99999994f);
}

// Rotate unit direction (u, v, w) by polar cos(theta) and azimuth phi in
// [0, 2 pi) (the hardware sine and cosine: a few ulp, ample for sampling).
__device__ __forceinline__ void rotate(float &u, float &v, float &w, float ct, float phi) {
    const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    const float cp = __cosf(phi), sp = __sinf(phi);
    if (fabsf(w) > 0.99999f) {
        u = st * cp;
        v = st * sp;
        w = (w > 0.0f ? 1.0f : -1.0f) * ct;
    } else {
        const float it = rsqrtf(1.0f - w * w), t = (1.0f - w * w) * it;
        const float sit = st * it;
        const float nu = u * ct + sit * (u * w * cp - v * sp);
        const float nv = v * ct + sit * (v * w * cp + u * sp);
        const float nw = w * ct - t * st * cp;
        u = nu;
        v = nv;
        w = nw;
    }
    const float n = rsqrtf(u * u + v * v + w * w);
    u *= n;
    v *= n;
    w *= n;
}

// Klein-Nishina by Kahn's method; returns cos(theta), updates E.
__device__ __forceinline__ float compton(Ranecu &rng, float &E) {
    const float k = E * (1.0f / 510.99895f), ik = 1.0f / k;
    const float b1 = (1.0f + 2.0f * k) / (9.0f + 2.0f * k);
    float eta, ct;
    for (int it = 0; it < 1000; it++) {
        const float r1 = ranecu(rng), r2 = ranecu(rng), r3 = ranecu(rng);
        if (r1 <= b1) {
            eta = 1.0f + 2.0f * k * r2;
            if (r3 <= 4.0f * (1.0f / eta - 1.0f / (eta * eta))) break;
        } else {
            eta = (1.0f + 2.0f * k) / (1.0f + 2.0f * k * r2);
            ct = 1.0f - (eta - 1.0f) * ik;
            if (r3 <= 0.5f * (ct * ct + 1.0f / eta)) break;
        }
    }
    ct = 1.0f - (eta - 1.0f) * ik;
    E = E / eta;
    return fminf(1.0f, fmaxf(-1.0f, ct));
}

// Thomson angular distribution (1 + cos^2) / 2 by rejection.
__device__ __forceinline__ float rayleigh(Ranecu &rng) {
    float c = 0.0f;
    for (int it = 0; it < 1000; it++) {
        c = 2.0f * ranecu(rng) - 1.0f;
        if (ranecu(rng) <= 0.5f * (1.0f + c * c)) break;
    }
    return c;
}

struct Photon {
    float x[3], d[3];
    float E;
    int nc, nr;
    bool alive;
    bool inside;
};

constexpr float TWO_PI = 6.283185307179586f;

// (E / 60 keV)^(-k) of the three processes from one log2 and three exp2.
__device__ __forceinline__ void cross_sections(const drr_mc_params_t &p, float E, float &fpe, float &fc, float &fr) {
    const float l = __log2f(E * (1.0f / 60.0f));
    fpe = exp2f(-p.k_pe * l);
    fc = exp2f(-p.k_compton * l);
    fr = exp2f(-p.k_rayleigh * l);
}
constexpr float E_MIN = 5.0f;  // keV: lower energies are absorbed locally
constexpr int BRICK = 16;      // voxels per brick edge (a power of two)

struct Bricks {
    const float4 *maj;  // [nbz][nby][nbx]: (max rho a_pe, max rho a_c, max rho a_r, 0)
    int nb[3];
};

}  // namespace

// One wave per brick: lane l reads voxels l, l + 64, ... of the brick in x-fastest
// order (rows of BRICK contiguous records), then the wave reduces.
__global__ __launch_bounds__(256) void k_mc_bricks(const uint2 *__restrict__ vox, int nx, int ny, int nz,
                                                   drr_mc_params_t p, Bricks b, int64_t n_bricks) {
    const int64_t bi = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (bi >= n_bricks) return;  // whole waves leave together
    const int bx = (int)(bi % b.nb[0]), by = (int)((bi / b.nb[0]) % b.nb[1]), bz = (int)(bi / ((int64_t)b.nb[0] * b.nb[1]));
    float mpe = 0.0f, mc = 0.0f, mr = 0.0f;
    for (int v = lane; v < BRICK * BRICK * BRICK; v += 64) {
        const int i = bx * BRICK + v % BRICK, j = by * BRICK + (v / BRICK) % BRICK, k = bz * BRICK + v / (BRICK * BRICK);
        if (i < nx && j < ny && k < nz) {
            const uint2 rv = vox[((int64_t)k * ny + j) * nx + i];
            const float rho = __uint_as_float(rv.x);
            const int m = (int)rv.y;
            mpe = fmaxf(mpe, rho * p.a_pe[m]);
            mc = fmaxf(mc, rho * p.a_compton[m]);
            mr = fmaxf(mr, rho * p.a_rayleigh[m]);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        mpe = fmaxf(mpe, __shfl_xor(mpe, o));
        mc = fmaxf(mc, __shfl_xor(mc, o));
        mr = fmaxf(mr, __shfl_xor(mr, o));
    }
    if (lane == 0) const_cast<float4 *>(b.maj)[bi] = make_float4(mpe, mc, mr, 0.0f);
}

// 7 waves per SIMD (72 VGPRs, one spilled): 3 % more histories/s than the
// compiler's own 74 VGPRs at 6 waves; 8 waves spill 11 and lose 3 %
__global__ __launch_bounds__(256, 7) void k_mc(const uint2 *__restrict__ vox, int nx, int ny, int nz, drr_mc_params_t p,
                                            Bricks bk, int64_t first, int64_t count, uint64_t seed, float *img,
                                            float *img2, unsigned long long *nsteps, unsigned long long *nbsteps) {
    __shared__ float s_e[DRR_MAX_BINS], s_cdf[DRR_MAX_BINS];
    __shared__ float s_a[3][DRR_MAX_MATERIALS];
    for (int i = threadIdx.x; i < p.n_bins; i += blockDim.x) {
        s_e[i] = p.energies_keV[i];
        s_cdf[i] = p.cdf[i];
    }
    for (int i = threadIdx.x; i < p.n_materials; i += blockDim.x) {
        s_a[0][i] = p.a_pe[i];
        s_a[1][i] = p.a_compton[i];
        s_a[2][i] = p.a_rayleigh[i];
    }
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int W = p.det_width, H = p.det_height;
    const int64_t npix = (int64_t)W * H;
    // pixel coordinates along u and v are (hit - origin) . u / |u|^2
    const float iuu = 1.0f / (p.detector_u[0] * p.detector_u[0] + p.detector_u[1] * p.detector_u[1] + p.detector_u[2] * p.detector_u[2]);
    const float ivv = 1.0f / (p.detector_v[0] * p.detector_v[0] + p.detector_v[1] * p.detector_v[1] + p.detector_v[2] * p.detector_v[2]);
    const float *A = p.ijk_from_world;
    unsigned long long steps = 0, bsteps = 0;
    Photon ph;
    ph.alive = false;
    Ranecu rng{1, 1};
    float fpe = 0.f, fc = 0.f, fr = 0.f;
    float t_exit = 0.f;
    for (;;) {
        if (!ph.alive && h >= count) break;
        {
            // a dead lane waits for a new history until 4 lanes of the wave are
            // dead or none is alive, so the long start of a history runs for
            // several lanes at once (each lane keeps its own sequence of
            // histories, so results do not change; measured against starting
            // at once: 4 lanes +3 %, 8 +1-2 %, 12 and 16 no change, 32 slower)
            const unsigned long long alive = __ballot(ph.alive);
            const int ndead = __popcll(__ballot(!ph.alive));
            if (!ph.alive && alive != 0ull && ndead < 4) continue;
        }
        if (!ph.alive) {
            rng = seed_history(seed, (uint64_t)(first + h));
            h += stride;
            // energy from the spectrum
            const float u = ranecu(rng);
            int lo = 0, hi = p.n_bins - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (s_cdf[mid] >= u) hi = mid; else lo = mid + 1;
            }
            ph.E = s_e[lo];
            // direction uniform in the cone about the axis
            ph.d[0] = p.cone_axis[0];
            ph.d[1] = p.cone_axis[1];
            ph.d[2] = p.cone_axis[2];
            const float ct = 1.0f - ranecu(rng) * (1.0f - p.cone_cos);
            rotate(ph.d[0], ph.d[1], ph.d[2], ct, TWO_PI * ranecu(rng));
            ph.x[0] = p.source_world[0];
            ph.x[1] = p.source_world[1];
            ph.x[2] = p.source_world[2];
            ph.nc = ph.nr = 0;
            ph.alive = true;
            // entry into the volume box [-0.5, n-0.5] in IJK
            float tmin = 0.0f, tmax = INFINITY;
            const float nn[3] = {(float)nx, (float)ny, (float)nz};
            bool hit = true;
            for (int a = 0; a < 3; a++) {
                const float q0 = A[4 * a] * ph.x[0] + A[4 * a + 1] * ph.x[1] + A[4 * a + 2] * ph.x[2] + A[4 * a + 3];
                const float dq = A[4 * a] * ph.d[0] + A[4 * a + 1] * ph.d[1] + A[4 * a + 2] * ph.d[2];
                if (dq != 0.0f) {
                    const float idq = 1.0f / dq;
                    const float t0 = (-0.5f - q0) * idq, t1 = (nn[a] - 0.5f - q0) * idq;
                    tmin = fmaxf(tmin, fminf(t0, t1));
                    tmax = fminf(tmax, fmaxf(t0, t1));
                } else if (q0 < -0.5f || q0 > nn[a] - 0.5f) {
                    hit = false;
                }
            }
            ph.inside = hit && tmin < tmax;
            if (ph.inside) {
                for (int a = 0; a < 3; a++) ph.x[a] += tmin * ph.d[a];
                t_exit = tmax - tmin;
                cross_sections(p, ph.E, fpe, fc, fr);
            }
        }
        if (ph.inside) {
            // one flight: optical depth tau spent brick by brick at each
            // brick's majorant, up to a tentative collision or the box exit
            ++steps;
            float tau = -__logf(ranecu(rng));
            int bi[3], sg[3];
            float tn[3], td[3];
            for (int a = 0; a < 3; a++) {
                const float qp = A[4 * a] * ph.x[0] + A[4 * a + 1] * ph.x[1] + A[4 * a + 2] * ph.x[2] + A[4 * a + 3] + 0.5f;
                const float dq = A[4 * a] * ph.d[0] + A[4 * a + 1] * ph.d[1] + A[4 * a + 2] * ph.d[2];
                bi[a] = min(max((int)floorf(qp * (1.0f / BRICK)), 0), bk.nb[a] - 1);
                const float idq = 1.0f / dq;  // one division per axis, not two
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
            float t = 0.0f, mub = 0.0f, s = 0.0f;
            bool coll = false;
            for (;;) {
                ++bsteps;
                const float4 mj = bk.maj[((int64_t)bi[2] * bk.nb[1] + bi[1]) * bk.nb[0] + bi[0]];
                mub = (mj.x * fpe + mj.y * fc) + mj.z * fr;
                const int a = (tn[0] <= tn[1] && tn[0] <= tn[2]) ? 0 : (tn[1] <= tn[2] ? 1 : 2);
                const float tnext = fminf(tn[a], t_exit);
                const float seg = tnext - t;
                if (mub > 0.0f) {
                    if (tau < mub * seg) {
                        s = t + tau / mub;
                        coll = true;
                        break;
                    }
                    tau -= mub * seg;
                }
                t = tnext;
                if (tnext >= t_exit) break;
                bi[a] += sg[a];
                if (bi[a] < 0 || bi[a] >= bk.nb[a]) break;
                tn[a] += td[a];
            }
            if (!coll) {
                for (int a = 0; a < 3; a++) ph.x[a] += t_exit * ph.d[a];
                ph.inside = false;  // left the box: fly to the detector below
            } else {
                for (int a = 0; a < 3; a++) ph.x[a] += s * ph.d[a];
                t_exit -= s;
                int ix[3];
                const int nn[3] = {nx, ny, nz};
                for (int a = 0; a < 3; a++) {
                    const float q = A[4 * a] * ph.x[0] + A[4 * a + 1] * ph.x[1] + A[4 * a + 2] * ph.x[2] + A[4 * a + 3];
                    ix[a] = min(max((int)floorf(q + 0.5f), 0), nn[a] - 1);
                }
                const int64_t vi = ((int64_t)ix[2] * ny + ix[1]) * nx + ix[0];
                const uint2 rv = vox[vi];  // (density bits, label)
                const float rho = __uint_as_float(rv.x);
                const int m = (int)rv.y;
                const float mpe = rho * s_a[0][m] * fpe;
                const float mc = rho * s_a[1][m] * fc;
                const float mr = rho * s_a[2][m] * fr;
                const float mt = mpe + mc + mr;
                if (ranecu(rng) * mub < mt) {
                    const float u = ranecu(rng) * mt;
                    float ct;
                    if (u < mpe) {
                        ph.alive = false;  // photoelectric absorption
                        continue;
                    } else if (u < mpe + mc) {
                        ct = compton(rng, ph.E);
                        ph.nc++;
                    } else {
                        ct = rayleigh(rng);
                        ph.nr++;
                    }
                    rotate(ph.d[0], ph.d[1], ph.d[2], ct, TWO_PI * ranecu(rng));
                    if (ph.E < E_MIN || ph.nc + ph.nr > p.max_interactions) {
                        ph.alive = false;
                        continue;
                    }
                    // new direction / energy: distance to the box exit and majorant
                    float tmax = INFINITY;
                    const float nnf[3] = {(float)nx, (float)ny, (float)nz};
                    for (int a = 0; a < 3; a++) {
                        const float q0 = A[4 * a] * ph.x[0] + A[4 * a + 1] * ph.x[1] + A[4 * a + 2] * ph.x[2] + A[4 * a + 3];
                        const float dq = A[4 * a] * ph.d[0] + A[4 * a + 1] * ph.d[1] + A[4 * a + 2] * ph.d[2];
                        if (dq != 0.0f) {
                            const float idq = 1.0f / dq;
                            tmax = fminf(tmax, fmaxf((-0.5f - q0) * idq, (nnf[a] - 0.5f - q0) * idq));
                        }
                    }
                    t_exit = fmaxf(tmax, 0.0f);
                    cross_sections(p, ph.E, fpe, fc, fr);
                }
            }
            continue;
        }
        // outside the box, straight to the detector plane
        ph.alive = false;
        const float dn = ph.d[0] * p.detector_normal[0] + ph.d[1] * p.detector_normal[1] + ph.d[2] * p.detector_normal[2];
        if (dn <= 0.0f) continue;
        const float t = ((p.detector_origin[0] - ph.x[0]) * p.detector_normal[0] +
                         (p.detector_origin[1] - ph.x[1]) * p.detector_normal[1] +
                         (p.detector_origin[2] - ph.x[2]) * p.detector_normal[2]) / dn;
        if (t < 0.0f) continue;
        float r[3];
        for (int a = 0; a < 3; a++) r[a] = ph.x[a] + t * ph.d[a] - p.detector_origin[a];
        const float fu = (r[0] * p.detector_u[0] + r[1] * p.detector_u[1] + r[2] * p.detector_u[2]) * iuu;
        const float fv = (r[0] * p.detector_v[0] + r[1] * p.detector_v[1] + r[2] * p.detector_v[2]) * ivv;
        if (!(fu >= 0.0f && fv >= 0.0f && fu < (float)W && fv < (float)H)) continue;
        const int iu = min((int)fu, W - 1), iv = min((int)fv, H - 1);
        const int cls = (ph.nc + ph.nr == 0) ? 0 : (ph.nc == 1 && ph.nr == 0) ? 1 : (ph.nr == 1 && ph.nc == 0) ? 2 : 3;
        const int64_t o = cls * npix + (int64_t)iv * W + iu;
        unsafeAtomicAdd(img + o, ph.E);
        if (img2) unsafeAtomicAdd(img2 + o, ph.E * ph.E);
    }
    if (nsteps) {
        for (int o = 32; o > 0; o >>= 1) steps += __shfl_xor(steps, o);
        if ((threadIdx.x & 63) == 0) atomicAdd(nsteps, steps);
    }
    if (nbsteps) {
        for (int o = 32; o > 0; o >>= 1) bsteps += __shfl_xor(bsteps, o);
        if ((threadIdx.x & 63) == 0) atomicAdd(nbsteps, bsteps);
    }
}

extern "C" int drr_mc_scatter(drr_ctx *ctx, const drr_volume *vol, const drr_mc_params_t *params,
                              int64_t first_history, int64_t n_histories, uint64_t seed, float *image,
                              float *image2, int64_t *n_steps, int64_t *n_brick_steps, void *stream) {
    if (!ctx || !vol || !params || !image) return drr_set_error(DRR_ERR_INVALID, "drr_mc_scatter: NULL argument");
    if (vol->ctx != ctx) return drr_set_error(DRR_ERR_INVALID, "drr_mc_scatter: volume of another context");
    const drr_mc_params_t &p = *params;
    if (p.n_materials < 1 || p.n_materials > DRR_MAX_MATERIALS || p.n_bins < 1 || p.n_bins > DRR_MAX_BINS)
        return drr_set_error(DRR_ERR_INVALID, "drr_mc_scatter: n_materials in [1,16] and n_bins in [1,256]");
    if (p.det_width < 1 || p.det_height < 1 || first_history < 0 || n_histories < 0)
        return drr_set_error(DRR_ERR_INVALID, "drr_mc_scatter: bad detector size or history range");
    if (!(p.cone_cos > -1.0f && p.cone_cos < 1.0f))
        return drr_set_error(DRR_ERR_INVALID, "drr_mc_scatter: cone_cos must be in (-1, 1)");
    if (!(p.cdf[p.n_bins - 1] >= 0.999f))
        return drr_set_error(DRR_ERR_INVALID, "drr_mc_scatter: cdf must end at 1");
    if (p.max_interactions < 1) return drr_set_error(DRR_ERR_INVALID, "drr_mc_scatter: max_interactions >= 1");
    if (ctx->have_scene && p.n_materials < ctx->scene.n_materials)
        return drr_set_error(DRR_ERR_INVALID, "drr_mc_scatter: tables must cover every label of the scene");
    if (n_histories == 0) return DRR_OK;
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != ctx->device) DRR_HIP_CHECK(hipSetDevice(ctx->device));
    drr_volume *v = const_cast<drr_volume *>(vol);
    int rc = drr_volume_ensure_linear(v);
    if (rc) return rc;
    const drr_mc_params_t &pp = p;
    int cus = 0;
    DRR_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    // Every lane takes histories h, h + stride, ...: a grid larger than what is
    // resident at once would leave its later blocks a full share of work to
    // start only when the first ones finish, so the grid is capped there.
    int per_cu = 0;
    DRR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mc, 256, 0));
    int64_t blocks = (n_histories + 255) / 256;
    const int64_t cap = (int64_t)cus * (per_cu > 0 ? per_cu : 1);
    if (blocks > cap) blocks = cap;
    hipStream_t st = (hipStream_t)stream;
    // the brick majorants of these cross sections, in stream-ordered memory
    Bricks bk{};
    bk.nb[0] = (v->nx + BRICK - 1) / BRICK;
    bk.nb[1] = (v->ny + BRICK - 1) / BRICK;
    bk.nb[2] = (v->nz + BRICK - 1) / BRICK;
    const int64_t n_bricks = (int64_t)bk.nb[0] * bk.nb[1] * bk.nb[2];
    float4 *maj = nullptr;
    hipError_t e = hipMallocAsync((void **)&maj, sizeof(float4) * n_bricks, st);
    if (e != hipSuccess) {
        if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
        return drr_hip_error(e, "hipMallocAsync(brick majorants)");
    }
    bk.maj = maj;
    (void)hipEventRecord(ctx->ev0, st);
    k_mc_bricks<<<(unsigned)((n_bricks + 3) / 4), 256, 0, st>>>(v->lin_vox, v->nx, v->ny, v->nz, pp, bk, n_bricks);
    k_mc<<<(unsigned)blocks, 256, 0, st>>>(v->lin_vox, v->nx, v->ny, v->nz, pp, bk, first_history,
                                         n_histories, seed, image, image2,
                                         reinterpret_cast<unsigned long long *>(n_steps),
                                         reinterpret_cast<unsigned long long *>(n_brick_steps));
    e = hipGetLastError();
    (void)hipEventRecord(ctx->ev1, st);
    ctx->timed = true;
    (void)hipFreeAsync(maj, st);
    if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
    if (e != hipSuccess) return drr_hip_error(e, "drr_mc_scatter launch");
    return DRR_OK;
}
