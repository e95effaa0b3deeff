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
// gradient from dL/dp_t = wt_t [ (sum_d w_d h_d) grad(rho) + rho grad(sum_d h_d w_d) ].
// Nothing is stored per sample.
#include <hip/hip_runtime.h>

#include "drr_internal.h"
#include "march.h"

using namespace drr;

struct AdjArgs {
    const float *gI;   // [H][W] or NULL
    const float *gP;   // [H][W] or NULL
    float sw[3];       // source in world
    float c[3];        // pose centre in world
    float *gD;         // [nz][ny][nx]
    float *gPose;      // [6] or NULL
};

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
    const int cx = clamp_cell(bx, vol.nx), cy = clamp_cell(by, vol.ny), cz = clamp_cell(bz, vol.nz);
#pragma unroll
    for (int d = 0; d < 8; d++) {
        if (acc[d] != 0.0f) {
            const int x = min(max(cx + (d & 1), 0), vol.nx - 1);
            const int y = min(max(cy + ((d >> 1) & 1), 0), vol.ny - 1);
            const int z = min(max(cz + (d >> 2), 0), vol.nz - 1);
            unsafeAtomicAdd(gD + ((size_t)z * vol.ny + y) * vol.nx + x, acc[d]);
        }
        acc[d] = 0.0f;
    }
}

template <int M, bool POSE>
__device__ __forceinline__ void adjoint_march(const Geo1 &g, const VolDev &vol, int udx, int vdx,
                                              float step, const float (&graw)[M], const AdjArgs &a,
                                              double (&gpose)[6]) {
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
    while (t < N && alpha < lim) {
        alpha += step;
        ++t;
    }
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
    float acc[8];
#pragma unroll
    for (int d = 0; d < 8; d++) acc[d] = 0.0f;
    float h[8];
    float hu = 0.0f;  // h of a uniform cell
    auto set_h = [&]() {
        if (c.uni) {
            hu = get_m<M>(graw, (int)c.m0);
        } else {
#pragma unroll
            for (int d = 0; d < 8; d++) h[d] = get_m<M>(graw, (int)corner_label(c.l0, c.l1, d));
        }
    };
    if (c.have) set_h();
    // per-ray pose sums in FP64: they add thousands of terms of both signs
    double S0[3] = {0.0, 0.0, 0.0}, S1[3] = {0.0, 0.0, 0.0};
    while (t < N && alpha <= r.mxv) {
        const float px = sx + alpha * rix - 1.0f;
        const float py = sy + alpha * riy - 1.0f;
        const float pz = sz + alpha * riz - 1.0f;
        const float bx = floorf(px), by = floorf(py), bz = floorf(pz);
        if (!c.have || bx != c.bx || by != c.by || bz != c.bz) {
            if (c.have) flush_cell(vol, a.gD, c.bx, c.by, c.bz, acc);
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
            for (int d = 0; d < 8; d++) sw += w[d] * h[d];
        }
        const float G = wt * sw;
#pragma unroll
        for (int d = 0; d < 8; d++) acc[d] += G * w[d];
        if (POSE) {
            const float D[8] = {c.d0.x, c.d0.y, c.d0.z, c.d0.w, c.d1.x, c.d1.y, c.d1.z, c.d1.w};
            float gr[3];
            grad_trilinear(D, fx, fy, fz, gr);
            const float rho = trilerp(c.d0, c.d1, w);
            float gh[3] = {0.f, 0.f, 0.f};
            if (!c.uni) grad_trilinear(h, fx, fy, fz, gh);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float gp = wt * (sw * gr[k] + rho * gh[k]);
                S0[k] += (double)gp;
                S1[k] += (double)alpha * (double)gp;
            }
        }
        alpha += step;
        ++t;
    }
    if (c.have) flush_cell(vol, a.gD, c.bx, c.by, c.bz, acc);
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
        const double q[3] = {(double)a.sw[0] - a.c[0], (double)a.sw[1] - a.c[1], (double)a.sw[2] - a.c[2]};
        const double rw[3] = {r.rx, r.ry, r.rz};
        gpose[0] -= V0[0];
        gpose[1] -= V0[1];
        gpose[2] -= V0[2];
        gpose[3] += (V0[1] * q[2] - V0[2] * q[1]) + (V1[1] * rw[2] - V1[2] * rw[1]);
        gpose[4] += (V0[2] * q[0] - V0[0] * q[2]) + (V1[2] * rw[0] - V1[0] * rw[2]);
        gpose[5] += (V0[0] * q[1] - V0[1] * q[0]) + (V1[0] * rw[1] - V1[1] * rw[0]);
    }
}

__device__ __forceinline__ int xcd_remap_adj(int bid, int nb) {
    const int xcd = bid & 7, idx = bid >> 3;
    const int per = nb >> 3, rem = nb & 7;
    return xcd * per + min(xcd, rem) + idx;
}

template <int M, bool ATTEN, bool POSE>
__global__ __launch_bounds__(256) void k_adjoint_v1(Geo1 g, VolDev vol, SpectrumDev sp, float step,
                                                    int air, AdjArgs a, int tiles16_x) {
    const int lb = xcd_remap_adj(blockIdx.x, gridDim.x);
    const int bx = lb % tiles16_x, by = lb / tiles16_x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int udx, vdx;
    lane_pixel(bx * 2 + (wave & 1), by * 2 + (wave >> 1), lane, udx, vdx);
    double gpose[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (udx < g.width && vdx < g.height) {
        const int64_t idx = (int64_t)vdx * g.width + udx;
        const float gi = a.gI ? a.gI[idx] : 0.0f;
        const float gp = a.gP ? a.gP[idx] : 0.0f;
        if (gi != 0.0f || gp != 0.0f) {
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
            adjoint_march<M, POSE>(g, vol, udx, vdx, step, graw, a, gpose);
        }
    }
    if (POSE) {
#pragma unroll
        for (int k = 0; k < 6; k++) {
            double v = gpose[k];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if (lane == 0 && v != 0.0) unsafeAtomicAdd(a.gPose + k, (float)v);
        }
    }
}

template <int M, bool ATTEN>
static void launch_adj(bool pose, dim3 gr, dim3 bl, hipStream_t st, const Geo1 &g, const VolDev &v,
                       const SpectrumDev &sp, float step, int air, const AdjArgs &a, int tx) {
    if (pose)
        k_adjoint_v1<M, ATTEN, true><<<gr, bl, 0, st>>>(g, v, sp, step, air, a, tx);
    else
        k_adjoint_v1<M, ATTEN, false><<<gr, bl, 0, st>>>(g, v, sp, step, air, a, tx);
}

extern "C" int drr_adjoint(drr_ctx *ctx, const drr_volume *vol, const drr_view_t *view,
                           const float *grad_intensity, const float *grad_photon_prob,
                           const float *source_world, const float *pose_center_world,
                           float *grad_density, float *grad_pose, void *stream) {
    const drr_volume *vols[1] = {vol};
    int rc = drr_check_view(ctx, vols, view);
    if (rc) return rc;
    if (view->n_volumes != 1) return drr_set_error(DRR_ERR_UNSUPPORTED, "drr_adjoint: one volume per view");
    if (!grad_density) return drr_set_error(DRR_ERR_INVALID, "drr_adjoint: grad_density is NULL");
    if (!grad_intensity && !grad_photon_prob)
        return drr_set_error(DRR_ERR_INVALID, "drr_adjoint: no output gradient given");
    if (grad_pose && (!source_world || !pose_center_world))
        return drr_set_error(DRR_ERR_INVALID, "drr_adjoint: pose gradient needs source_world and pose_center_world");
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
    for (int i = 0; i < 3; i++) {
        a.sw[i] = source_world ? source_world[i] : 0.0f;
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
#define DRR_CASE(MM)                                                                          \
    case MM:                                                                                  \
        if (sc.attenuate_outside)                                                             \
            launch_adj<MM, true>(pose, grid, block, st, g, v, sp, sc.step, sc.air_index, a, tx);  \
        else                                                                                  \
            launch_adj<MM, false>(pose, grid, block, st, g, v, sp, sc.step, sc.air_index, a, tx); \
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
