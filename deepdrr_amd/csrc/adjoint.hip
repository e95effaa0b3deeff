// adjoint.hip -- reverse-mode derivative of the single-volume forward
// projector (drr_project) with respect to the density volume and a rigid
// 6-DoF perturbation of the volume pose.
//
// deepdrr itself is not differentiable (README.md:47); this kernel is the
// exact adjoint of the forward march in march.h (project_kernel.cu:398-646
// semantics): for every ray
//   ad[m]  = step/10 * sum_t wt_t rho(p_t) seg_m(p_t)
//   I      = sum_b E_b pdf_b exp(-sum_m ad[m] mu[b][m]),  P likewise without E_b
// pass 1 re-runs the forward march for ad[], pass 2 re-marches the same
// samples and scatters
//   dL/dD_corner += wt_t * (sum_d w_d h_d) * w_corner,   h_d = dL/dad_raw[label_d]
// with one atomic flush per cell change (per-corner sums are kept in
// registers while the ray stays in the cell), and accumulates the pose
// gradient from dL/dp_t = wt_t [ (sum_d w_d h_d) grad(rho) + rho grad(sum_d h_d w_d) ]
// (the second term vanishes in a cell whose labels agree).  Nothing is stored
// per sample.  drr_adjoint_jobs runs many views in one persistent launch.
#include <hip/hip_runtime.h>

#include "drr_internal.h"
#include "jobs.h"
#include "march.h"

using namespace drr;

// d/d(fx, fy, fz) of sum_d v_d w_d for corner values v (corner d = dx + 2dy + 4dz).
__device__ __forceinline__ void grad_trilinear(const float v[8], float fx, float fy, float fz,
                                               float out[3]) {
    const float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
    out[0] = (v[1] - v[0]) * (gy * gz) + (v[3] - v[2]) * (fy * gz) + (v[5] - v[4]) * (gy * fz) +
             (v[7] - v[6]) * (fy * fz);
    out[1] = (v[2] - v[0]) * (gx * gz) + (v[3] - v[1]) * (fx * gz) + (v[6] - v[4]) * (gx * fz) +
             (v[7] - v[5]) * (fx * fz);
    out[2] = (v[4] - v[0]) * (gx * gy) + (v[5] - v[1]) * (fx * gy) + (v[6] - v[2]) * (gx * fy) +
             (v[7] - v[3]) * (fx * fy);
}

__device__ __forceinline__ void flush_cell(const VolDev &vol, float *gD, float bx, float by, float bz,
                                           float acc[8]) {
    // cell (cx, cy, cz) in [-1, n-1]^3; its corners clamp to [0, n-1] (voxels < 2^31)
    const int cx = clamp_cell(bx, vol.nx), cy = clamp_cell(by, vol.ny), cz = clamp_cell(bz, vol.nz);
    const uint32_t x0 = max(cx, 0), x1 = min(cx + 1, vol.nx - 1);
    const uint32_t y0 = max(cy, 0), y1 = min(cy + 1, vol.ny - 1);
    const uint32_t z0 = max(cz, 0), z1 = min(cz + 1, vol.nz - 1);
    const uint32_t nx = vol.nx, ny = vol.ny;
    const uint32_t row[4] = {(z0 * ny + y0) * nx, (z0 * ny + y1) * nx, (z1 * ny + y0) * nx, (z1 * ny + y1) * nx};
#pragma unroll
    for (int d = 0; d < 8; d++) {
        if (acc[d] != 0.0f) unsafeAtomicAdd(gD + (row[d >> 1] + ((d & 1) ? x1 : x0)), acc[d]);
        acc[d] = 0.0f;
    }
}

// Pass 2 of one ray: scatter density gradients, accumulate pose gradients.
template <int M, bool POSE>
__device__ __forceinline__ void adjoint_march(const Geo1 &g, const VolDev &vol, int udx, int vdx,
                                              float step, const float (&graw)[M], const float sw_world[3],
                                              const float center[3], float *gD, double (&gpose)[6]) {
    RayV1 r;
    march_setup_v1(g, udx, vdx, step, r);
    if (!r.trace) return;
    const int N = r.N;
    float alpha = r.minA;
    int t = 0;
    const float sx = g.s[0], sy = g.s[1], sz = g.s[2];
    const float rix = r.ri[0], riy = r.ri[1], riz = r.ri[2];
    const float rdom = fmaxf(fabsf(rix), fmaxf(fabsf(riy), fabsf(riz)));
    const float lim = r.mnv - (1.0f / rdom + 2.0f * step);
    skip_steps(N, lim, step, alpha, t);
    Cell c;
    while (t < N && alpha < r.mnv) {
        const float px = sx + alpha * rix - 1.0f;
        const float py = sy + alpha * riy - 1.0f;
        const float pz = sz + alpha * riz - 1.0f;
        const float bx = floorf(px), by = floorf(py), bz = floorf(pz);
        if (!c.have || bx != c.bx || by != c.by || bz != c.bz) c.load(vol, px, py, pz, bx, by, bz);
        alpha += step;
        ++t;
    }
    float acc[8];   // sum over the cell's samples of G w_d: density gradient of corner d
    float acc1[8];  // sum of alpha G w_d (pose only)
#pragma unroll
    for (int d = 0; d < 8; d++) {
        acc[d] = 0.0f;
        acc1[d] = 0.0f;
    }
    float h[8];
    float hu = 0.0f;  // h of a cell whose labels agree
    auto set_h = [&]() {
        if (c.uni) {
            hu = get_m<M>(graw, (int)c.m0);
        } else {
#pragma unroll
            for (int d = 0; d < 8; d++) h[d] = get_m<M>(graw, (int)corner_label(c.l0, c.l1, d));
        }
    };
    if (c.have) set_h();
    // Pose sums S0 = sum_t dL/dp_t, S1 = sum_t alpha_t dL/dp_t: FP32 within a
    // cell, FP64 across cells (thousands of terms of both signs).  Within a
    // cell, dL/dp_t = G_t grad(rho)(p_t) [+ wt rho_t grad(h)(p_t) if the labels
    // differ], and grad(rho) along x is sum_yz (D[1yz] - D[0yz]) w_yz, so
    // sum_t G_t grad_x(rho) = sum_yz (D[1yz] - D[0yz]) (acc[0yz] + acc[1yz])
    // (w_0yz + w_1yz = w_yz): the G term needs no per-sample gradient.
    double S0[3] = {0.0, 0.0, 0.0}, S1[3] = {0.0, 0.0, 0.0};
    float c0[3] = {0.f, 0.f, 0.f}, c1[3] = {0.f, 0.f, 0.f};  // label-gradient term of mixed cells
    auto flush_pose = [&]() {
        const float D[8] = {c.d0.x, c.d0.y, c.d0.z, c.d0.w, c.d1.x, c.d1.y, c.d1.z, c.d1.w};
        float g0[3] = {0.f, 0.f, 0.f}, g1[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int b = 1 << k;
#pragma unroll
            for (int d = 0; d < 8; d++) {
                if (d & b) continue;
                const float dD = D[d | b] - D[d];
                g0[k] = fmaf(dD, acc[d] + acc[d | b], g0[k]);
                g1[k] = fmaf(dD, acc1[d] + acc1[d | b], g1[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            S0[k] += (double)g0[k] + (double)c0[k];
            S1[k] += (double)g1[k] + (double)c1[k];
            c0[k] = 0.0f;
            c1[k] = 0.0f;
        }
#pragma unroll
        for (int d = 0; d < 8; d++) acc1[d] = 0.0f;
    };
    while (t < N && alpha <= r.mxv) {
        const float px = sx + alpha * rix - 1.0f;
        const float py = sy + alpha * riy - 1.0f;
        const float pz = sz + alpha * riz - 1.0f;
        const float bx = floorf(px), by = floorf(py), bz = floorf(pz);
        if (!c.have || bx != c.bx || by != c.by || bz != c.bz) {
            if (c.have) {
                if (POSE) flush_pose();
                flush_cell(vol, gD, c.bx, c.by, c.bz, acc);
            }
            c.load(vol, px, py, pz, bx, by, bz);
            set_h();
        }
        const float fx = px - bx, fy = py - by, fz = pz - bz;
        float w[8];
        weights8(fx, fy, fz, w);
        const float wt = (0 == t || N - 1 == t) ? 0.5f : 1.0f;
        float sw;
        if (c.uni) {
            const float ssum = ((((((w[0] + w[1]) + w[2]) + w[3]) + w[4]) + w[5]) + w[6]) + w[7];
            sw = hu * ssum;
        } else {
            sw = 0.0f;
#pragma unroll
            for (int d = 0; d < 8; d++) sw = fmaf(w[d], h[d], sw);
        }
        const float G = wt * sw;
#pragma unroll
        for (int d = 0; d < 8; d++) acc[d] = fmaf(G, w[d], acc[d]);
        if (POSE) {
            const float aG = alpha * G;
#pragma unroll
            for (int d = 0; d < 8; d++) acc1[d] = fmaf(aG, w[d], acc1[d]);
            if (!c.uni) {
                const float rho = trilerp(c.d0, c.d1, w);
                float gh[3];
                grad_trilinear(h, fx, fy, fz, gh);
                const float wr = wt * rho;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    c0[k] = fmaf(wr, gh[k], c0[k]);
                    c1[k] = fmaf(alpha * wr, gh[k], c1[k]);
                }
            }
        }
        alpha += step;
        ++t;
    }
    if (c.have) {
        if (POSE) flush_pose();
        flush_cell(vol, gD, c.bx, c.by, c.bz, acc);
    }
    if (POSE) {
        // p = A x + b - 1 with x = s_w + alpha r_w; perturbed x' = R(w)^T (x - c - t) + c:
        // dp/dt = -A, dp/dw = A [x - c]_x  ->  dL/dt = -A^T S0,
        // dL/dw = (A^T S0) x (s_w - c) + (A^T S1) x r_w
        const float *A = g.A;
        double V0[3], V1[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            V0[j] = (double)A[j] * S0[0] + (double)A[4 + j] * S0[1] + (double)A[8 + j] * S0[2];
            V1[j] = (double)A[j] * S1[0] + (double)A[4 + j] * S1[1] + (double)A[8 + j] * S1[2];
        }
        const double q[3] = {(double)sw_world[0] - center[0], (double)sw_world[1] - center[1],
                             (double)sw_world[2] - center[2]};
        const double rw[3] = {r.rx, r.ry, r.rz};
        gpose[0] -= V0[0];
        gpose[1] -= V0[1];
        gpose[2] -= V0[2];
        gpose[3] += (V0[1] * q[2] - V0[2] * q[1]) + (V1[1] * rw[2] - V1[2] * rw[1]);
        gpose[4] += (V0[2] * q[0] - V0[0] * q[2]) + (V1[2] * rw[0] - V1[0] * rw[2]);
        gpose[5] += (V0[0] * q[1] - V0[1] * q[0]) + (V1[0] * rw[1] - V1[1] * rw[0]);
    }
}

// Both passes for one pixel with output gradients gi (intensity), gp (photon_prob).
template <int M, bool ATTEN, bool POSE>
__device__ __forceinline__ void adjoint_ray(const Geo1 &g, const VolDev &vol, const SpectrumDev &sp, float step,
                                            int air, int udx, int vdx, float gi, float gp, const float sw[3],
                                            const float center[3], float *gD, double (&gpose)[6]) {
    if (gi == 0.0f && gp == 0.0f) return;
    float ad[M];
    march_v1<M, ATTEN>(g, vol, udx, vdx, step, air, ad);
    float gad[M];
#pragma unroll
    for (int m = 0; m < M; m++) gad[m] = 0.0f;
    for (int b = 0; b < sp.n_bins; b++) {
        float e = 0.0f;
#pragma unroll
        for (int m = 0; m < M; m++) e += ad[m] * sp.mu[b * M + m];
        const float coef = (gi * sp.energies[b] + gp) * sp.pdf[b] * expf(-e);
#pragma unroll
        for (int m = 0; m < M; m++) gad[m] -= coef * sp.mu[b * M + m];
    }
    float graw[M];
#pragma unroll
    for (int m = 0; m < M; m++) graw[m] = gad[m] * step / 10.0f;
    adjoint_march<M, POSE>(g, vol, udx, vdx, step, graw, sw, center, gD, gpose);
}

__device__ __forceinline__ void flush_pose_wave(double (&gpose)[6], float *gPose) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        double v = gpose[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0 && v != 0.0) unsafeAtomicAdd(gPose + k, (float)v);
        gpose[k] = 0.0;
    }
}

struct AdjArgs {
    const float *gI;  // [H][W] (or job images at out_offset), may be NULL
    const float *gP;  // may be NULL
    float c[3];       // pose centre in world
    float *gD;        // [nz][ny][nx]
    float *gPose;     // [6] or NULL
};

__device__ __forceinline__ int xcd_remap_adj(int bid, int nb) {
    const int xcd = bid & 7, idx = bid >> 3;
    const int per = nb >> 3, rem = nb & 7;
    return xcd * per + min(xcd, rem) + idx;
}

template <int M, bool ATTEN, bool POSE>
__global__ __launch_bounds__(256) void k_adjoint_v1(Geo1 g, VolDev vol, SpectrumDev sp, float step, int air,
                                                    AdjArgs a, float swx, float swy, float swz, int tiles16_x) {
    const int lb = xcd_remap_adj(blockIdx.x, gridDim.x);
    const int bx = lb % tiles16_x, by = lb / tiles16_x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int udx, vdx;
    lane_pixel(bx * 2 + (wave & 1), by * 2 + (wave >> 1), lane, udx, vdx);
    double gpose[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (udx < g.width && vdx < g.height) {
        const int64_t idx = (int64_t)vdx * g.width + udx;
        const float sw[3] = {swx, swy, swz};
        adjoint_ray<M, ATTEN, POSE>(g, vol, sp, step, air, udx, vdx, a.gI ? a.gI[idx] : 0.0f,
                                    a.gP ? a.gP[idx] : 0.0f, sw, a.c, a.gD, gpose);
    }
    if (POSE) flush_pose_wave(gpose, a.gPose);
}

// Wave sum of the pose gradient into an FP64 accumulator (one atomic per wave).
__device__ __forceinline__ void flush_pose_wave64(const double (&gpose)[6], double *acc) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        double v = gpose[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0 && v != 0.0) unsafeAtomicAdd(acc + k, v);
    }
}

__global__ void k_pose_zero(double *acc) {
    if (threadIdx.x < 6) acc[threadIdx.x] = 0.0;
}

__global__ void k_pose_finish(const double *acc, float *gPose) {
    if (threadIdx.x < 6) gPose[threadIdx.x] += (float)acc[threadIdx.x];
}

template <int M, bool ATTEN, bool POSE>
__global__ __launch_bounds__(256) void k_adjoint_jobs(const JobDev *__restrict__ jobs, const int2 *__restrict__ items,
                                                      const int *__restrict__ seg, int *counters,
                                                      const float *__restrict__ sw_jobs, VolDev vol, SpectrumDev sp,
                                                      float step, int air, AdjArgs a, double *pose64) {
    for_each_tile<8>(jobs, items, seg, counters, [&](int j, const JobDev &J, int udx, int vdx, bool valid) {
        double gpose[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (valid) {
            const int64_t o = J.out_offset + (int64_t)vdx * J.g.width + udx;
            const float sw[3] = {sw_jobs[3 * j], sw_jobs[3 * j + 1], sw_jobs[3 * j + 2]};
            adjoint_ray<M, ATTEN, POSE>(J.g, vol, sp, step, air, udx, vdx, a.gI ? a.gI[o] : 0.0f,
                                        a.gP ? a.gP[o] : 0.0f, sw, a.c, a.gD, gpose);
        }
        // per tile, so the FP64 sums are not held in registers across tiles
        if (POSE) flush_pose_wave64(gpose, pose64);
    });
}

template <int M, bool ATTEN>
static void launch_adj(bool pose, dim3 gr, dim3 bl, hipStream_t st, const Geo1 &g, const VolDev &v,
                       const SpectrumDev &sp, float step, int air, const AdjArgs &a, const float *sw, int tx) {
    if (pose)
        k_adjoint_v1<M, ATTEN, true><<<gr, bl, 0, st>>>(g, v, sp, step, air, a, sw[0], sw[1], sw[2], tx);
    else
        k_adjoint_v1<M, ATTEN, false><<<gr, bl, 0, st>>>(g, v, sp, step, air, a, sw[0], sw[1], sw[2], tx);
}

template <int M, bool ATTEN>
static void launch_adj_jobs(bool pose, const JobTable &t, hipStream_t st, const float *sw_jobs, const VolDev &v,
                            const SpectrumDev &sp, float step, int air, const AdjArgs &a, double *pose64) {
    dim3 gr(t.grid), bl(t.block);
    if (pose)
        k_adjoint_jobs<M, ATTEN, true><<<gr, bl, 0, st>>>(t.d_jobs, t.d_items, t.d_seg, t.d_counters, sw_jobs, v,
                                                          sp, step, air, a, pose64);
    else
        k_adjoint_jobs<M, ATTEN, false><<<gr, bl, 0, st>>>(t.d_jobs, t.d_items, t.d_seg, t.d_counters, sw_jobs, v,
                                                           sp, step, air, a, pose64);
}

static int check_adj(drr_ctx *ctx, const drr_volume *vol, const float *gI, const float *gP, const float *gD,
                     const float *gPose, const float *sw, const float *c) {
    if (!ctx || !vol) return drr_set_error(DRR_ERR_INVALID, "drr_adjoint: NULL context or volume");
    if (vol->ctx != ctx) return drr_set_error(DRR_ERR_INVALID, "drr_adjoint: volume of another context");
    if (!ctx->have_scene || ctx->n_bins < 1) return drr_set_error(DRR_ERR_STATE, "drr_adjoint: scene/spectrum not set");
    if (!gD) return drr_set_error(DRR_ERR_INVALID, "drr_adjoint: grad_density is NULL");
    if (!gI && !gP) return drr_set_error(DRR_ERR_INVALID, "drr_adjoint: no output gradient given");
    if (gPose && (!sw || !c))
        return drr_set_error(DRR_ERR_INVALID, "drr_adjoint: pose gradient needs source_world and pose_center_world");
    return DRR_OK;
}

extern "C" int drr_adjoint(drr_ctx *ctx, const drr_volume *vol, const drr_view_t *view,
                           const float *grad_intensity, const float *grad_photon_prob,
                           const float *source_world, const float *pose_center_world,
                           float *grad_density, float *grad_pose, void *stream) {
    const drr_volume *vols[1] = {vol};
    int rc = drr_check_view(ctx, vols, view);
    if (rc) return rc;
    if (view->n_volumes != 1) return drr_set_error(DRR_ERR_UNSUPPORTED, "drr_adjoint: one volume per view");
    rc = check_adj(ctx, vol, grad_intensity, grad_photon_prob, grad_density, grad_pose, source_world,
                   pose_center_world);
    if (rc) return rc;
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != ctx->device) DRR_HIP_CHECK(hipSetDevice(ctx->device));
    Geo1 g;
    view_to_geo1(*view, g);
    const VolDev v = vol_to_dev(vol);
    SpectrumDev sp{ctx->d_energies, ctx->d_pdf, ctx->d_mu, ctx->n_bins};
    AdjArgs a{};
    a.gI = grad_intensity;
    a.gP = grad_photon_prob;
    float sw[3] = {0.f, 0.f, 0.f};
    for (int i = 0; i < 3; i++) {
        sw[i] = source_world ? source_world[i] : 0.0f;
        a.c[i] = pose_center_world ? pose_center_world[i] : 0.0f;
    }
    a.gD = grad_density;
    a.gPose = grad_pose;
    const drr_scene_t &sc = ctx->scene;
    const int tx = (view->width + 15) / 16, ty = (view->height + 15) / 16;
    dim3 grid((unsigned)(tx * ty)), block(256);
    hipStream_t st = (hipStream_t)stream;
    const bool pose = grad_pose != nullptr;
    (void)hipEventRecord(ctx->ev0, st);
    switch (sc.n_materials) {
#define DRR_CASE(MM)                                                                              \
    case MM:                                                                                      \
        if (sc.attenuate_outside)                                                                 \
            launch_adj<MM, true>(pose, grid, block, st, g, v, sp, sc.step, sc.air_index, a, sw, tx);  \
        else                                                                                      \
            launch_adj<MM, false>(pose, grid, block, st, g, v, sp, sc.step, sc.air_index, a, sw, tx); \
        break;
        DRR_CASE(1) DRR_CASE(2) DRR_CASE(3) DRR_CASE(4) DRR_CASE(5) DRR_CASE(6) DRR_CASE(7) DRR_CASE(8)
#undef DRR_CASE
    }
    hipError_t e = hipGetLastError();
    (void)hipEventRecord(ctx->ev1, st);
    ctx->timed = true;
    if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
    if (e != hipSuccess) return drr_hip_error(e, "drr_adjoint launch");
    return DRR_OK;
}

extern "C" int drr_adjoint_jobs(drr_ctx *ctx, const drr_volume *vol, const drr_job_t *jobs, int n_jobs,
                                const float *grad_intensity, const float *grad_photon_prob,
                                const float *source_world, const float *pose_center_world,
                                float *grad_density, float *grad_pose, void *stream) {
    int rc = check_adj(ctx, vol, grad_intensity, grad_photon_prob, grad_density, grad_pose, source_world,
                       pose_center_world);
    if (rc) return rc;
    if (!jobs || n_jobs < 1) return drr_set_error(DRR_ERR_INVALID, "drr_adjoint_jobs: no jobs");
    for (int j = 0; j < n_jobs; j++)
        if (jobs[j].volume_index != 0)
            return drr_set_error(DRR_ERR_INVALID, "drr_adjoint_jobs: every job must use volume 0");
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != ctx->device) DRR_HIP_CHECK(hipSetDevice(ctx->device));
    const drr_volume *vols[1] = {vol};
    JobTable t;
    const drr_batch_tuning_t tune{8, 8, 4, 8};
    rc = drr_jobs_build(ctx, vols, 1, jobs, n_jobs, tune, t);
    if (rc != DRR_OK) return rc;
    // scratch: FP64 pose accumulator [6], then the job sources [3 * n_jobs]
    double *d_pose64 = nullptr;
    hipError_t e = hipMalloc(&d_pose64, sizeof(double) * 6 + sizeof(float) * 3 * n_jobs);
    float *d_sw = e == hipSuccess ? reinterpret_cast<float *>(d_pose64 + 6) : nullptr;
    if (e == hipSuccess && source_world)
        e = hipMemcpy(d_sw, source_world, sizeof(float) * 3 * n_jobs, hipMemcpyHostToDevice);
    else if (e == hipSuccess)
        e = hipMemset(d_sw, 0, sizeof(float) * 3 * n_jobs);
    if (e != hipSuccess) {
        drr_jobs_free(t);
        (void)hipFree(d_pose64);
        return drr_hip_error(e, "drr_adjoint_jobs: scratch");
    }
    const VolDev v = vol_to_dev(vol);
    SpectrumDev sp{ctx->d_energies, ctx->d_pdf, ctx->d_mu, ctx->n_bins};
    AdjArgs a{};
    a.gI = grad_intensity;
    a.gP = grad_photon_prob;
    for (int i = 0; i < 3; i++) a.c[i] = pose_center_world ? pose_center_world[i] : 0.0f;
    a.gD = grad_density;
    a.gPose = grad_pose;
    const drr_scene_t &sc = ctx->scene;
    hipStream_t st = (hipStream_t)stream;
    const bool pose = grad_pose != nullptr;
    (void)hipEventRecord(ctx->ev0, st);
    k_jobs_reset<<<1, 64, 0, st>>>(t.d_counters);
    if (pose) k_pose_zero<<<1, 64, 0, st>>>(d_pose64);
    switch (sc.n_materials) {
#define DRR_CASE(MM)                                                                                  \
    case MM:                                                                                          \
        if (sc.attenuate_outside)                                                                     \
            launch_adj_jobs<MM, true>(pose, t, st, d_sw, v, sp, sc.step, sc.air_index, a, d_pose64);  \
        else                                                                                          \
            launch_adj_jobs<MM, false>(pose, t, st, d_sw, v, sp, sc.step, sc.air_index, a, d_pose64); \
        break;
        DRR_CASE(1) DRR_CASE(2) DRR_CASE(3) DRR_CASE(4) DRR_CASE(5) DRR_CASE(6) DRR_CASE(7) DRR_CASE(8)
#undef DRR_CASE
    }
    if (pose) k_pose_finish<<<1, 64, 0, st>>>(d_pose64, grad_pose);
    e = hipGetLastError();
    (void)hipEventRecord(ctx->ev1, st);
    ctx->timed = true;
    hipError_t e2 = hipStreamSynchronize(st);
    drr_jobs_free(t);
    (void)hipFree(d_pose64);
    if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
    if (e != hipSuccess) return drr_hip_error(e, "drr_adjoint_jobs launch");
    if (e2 != hipSuccess) return drr_hip_error(e2, "drr_adjoint_jobs");
    return DRR_OK;
}
