// forward.hip -- cell-expansion of volumes and the one-view-per-launch
// forward projector (drr_project), the MI355X counterpart of deepdrr's
// projectKernel launch (deepdrr/projector/projector.py:709-800,
// deepdrr/projector/project_kernel.cu:135-650).
#include <hip/hip_runtime.h>

#include "drr_internal.h"
#include "march.h"

using namespace drr;

// ---- volume cell expansion --------------------------------------------------

// One thread per cell (i, j, k) in [-1, n-1]^3, i fastest: writes its 40-byte
// record (drr_internal.h): the 8 clamped corner densities, the corner labels
// as nibbles and their common label or ~0 (labels are < 16), as five 8-byte
// stores.  Consecutive threads write consecutive records, so the stores are
// fully coalesced.
__global__ __launch_bounds__(256) void k_expand(const float *__restrict__ density,
                                                const uint8_t *__restrict__ labels, int nx, int ny,
                                                int nz, uint32_t *__restrict__ cells) {
    const int64_t ncx = nx + 1, ncy = ny + 1, ncz = nz + 1;
    const int64_t ncell = ncx * ncy * ncz;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(c % ncx) - 1;
        const int j = (int)((c / ncx) % ncy) - 1;
        const int k = (int)(c / (ncx * ncy)) - 1;
        float d[8];
        uint32_t l[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int x = min(max(i + (q & 1), 0), nx - 1);
            const int y = min(max(j + ((q >> 1) & 1), 0), ny - 1);
            const int z = min(max(k + (q >> 2), 0), nz - 1);
            const int64_t idx = ((int64_t)z * ny + y) * nx + x;
            d[q] = density[idx];
            l[q] = labels[idx];
        }
        uint32_t packed = 0;
        bool same = true;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            packed |= l[q] << (4 * q);
            same = same && (l[q] == l[0]);
        }
        uint2 *rec = reinterpret_cast<uint2 *>(cells + DRR_CELL_WORDS * c);  // 8-byte aligned
#pragma unroll
        for (int q = 0; q < 4; q++) rec[q] = make_uint2(__float_as_uint(d[2 * q]), __float_as_uint(d[2 * q + 1]));
        rec[4] = make_uint2(packed, same ? l[0] : 0xffffffffu);
    }
}

int drr_launch_expand(const float *density, const uint8_t *labels, int nx, int ny, int nz,
                      uint32_t *cells, hipStream_t st) {
    const int64_t ncell = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    int64_t blocks = (ncell + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    k_expand<<<(unsigned)blocks, 256, 0, st>>>(density, labels, nx, ny, nz, cells);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return drr_hip_error(e, "k_expand launch");
    return DRR_OK;
}

// ---- plain forward kernels ----------------------------------------------------

// Block b takes 16x16 pixel tile b in row-major order.  Blocks are dispatched
// round-robin over the 8 XCDs, so every XCD works on every band of detector
// rows: ray lengths (and so the work) vary strongly between bands, and giving
// each XCD a contiguous band of its own (for L2 locality) made the XCDs with
// the central bands finish last -- 4-8% slower at 1536^2, measured.

struct OutDev {
    float *intensity;
    float *photon_prob;
    float *area_density;  // [M][H][W] or NULL
    float *solid_angle;   // [H][W] or NULL
};

template <int M, bool ATTEN, bool FAST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MarchOcc<M>::waves))) void k_project_v1(Geo1 g, VolDev vol, SpectrumDev sp, float step,
                                                    int air, OutDev out, int tiles16_x) {
    const int lb = blockIdx.x;
    const int bx = lb % tiles16_x, by = lb / tiles16_x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int udx, vdx;
    lane_pixel(bx * 2 + (wave & 1), by * 2 + (wave >> 1), lane, udx, vdx);
    if (udx >= g.width || vdx >= g.height) return;
    float ad[M];
    march_v1<M, ATTEN, FAST>(g, vol, udx, vdx, step, air, ad);
    float I, P;
    spectral<M>(sp, ad, I, P);
    const int64_t npix = (int64_t)g.width * g.height;
    const int64_t idx = (int64_t)vdx * g.width + udx;
    out.intensity[idx] = I;
    out.photon_prob[idx] = P;
    if (out.area_density) {
#pragma unroll
        for (int m = 0; m < M; m++) out.area_density[m * npix + idx] = ad[m];
    }
    if (out.solid_angle) out.solid_angle[idx] = solid_angle(g.W, udx, vdx);
}

template <int V, int M>
__global__ __launch_bounds__(256) void k_project_multi(ViewDev vd, SpectrumDev sp, float step,
                                                       int atten, int air, OutDev out,
                                                       int tiles16_x) {
    const int lb = blockIdx.x;
    const int bx = lb % tiles16_x, by = lb / tiles16_x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int udx, vdx;
    lane_pixel(bx * 2 + (wave & 1), by * 2 + (wave >> 1), lane, udx, vdx);
    if (udx >= vd.width || vdx >= vd.height) return;
    float ad[M];
    march_multi<V, M>(vd, udx, vdx, step, atten, air, ad);
    float I, P;
    spectral<M>(sp, ad, I, P);
    const int64_t npix = (int64_t)vd.width * vd.height;
    const int64_t idx = (int64_t)vdx * vd.width + udx;
    out.intensity[idx] = I;
    out.photon_prob[idx] = P;
    if (out.area_density) {
#pragma unroll
        for (int m = 0; m < M; m++) out.area_density[m * npix + idx] = ad[m];
    }
    if (out.solid_angle) out.solid_angle[idx] = solid_angle(vd.W, udx, vdx);
}

template <int M>
static void launch_v1(bool atten, bool fast, dim3 g, dim3 b, hipStream_t st, const Geo1 &geo, const VolDev &vol,
                      const SpectrumDev &sp, float step, int air, const OutDev &o, int tx) {
    if (fast) {
        if (atten)
            k_project_v1<M, true, true><<<g, b, 0, st>>>(geo, vol, sp, step, air, o, tx);
        else
            k_project_v1<M, false, true><<<g, b, 0, st>>>(geo, vol, sp, step, air, o, tx);
    } else {
        if (atten)
            k_project_v1<M, true, false><<<g, b, 0, st>>>(geo, vol, sp, step, air, o, tx);
        else
            k_project_v1<M, false, false><<<g, b, 0, st>>>(geo, vol, sp, step, air, o, tx);
    }
}

template <int V>
static void launch_multi(int M, dim3 g, dim3 b, hipStream_t st, const ViewDev &vd,
                         const SpectrumDev &sp, float step, int atten, int air, const OutDev &o,
                         int tx) {
    switch (M) {
#define DRR_CASE(MM) \
    case MM: k_project_multi<V, MM><<<g, b, 0, st>>>(vd, sp, step, atten, air, o, tx); break;
        DRR_CASE(1) DRR_CASE(2) DRR_CASE(3) DRR_CASE(4) DRR_CASE(5) DRR_CASE(6) DRR_CASE(7) DRR_CASE(8) \
        DRR_CASE(9) DRR_CASE(10) DRR_CASE(11) DRR_CASE(12) DRR_CASE(13) DRR_CASE(14) DRR_CASE(15) DRR_CASE(16)
#undef DRR_CASE
    }
}

extern "C" int drr_project(drr_ctx *ctx, const drr_volume *const *vols, const drr_view_t *view,
                           float *intensity, float *photon_prob, float *area_density,
                           float *solid_angle_out, void *stream) {
    int rc = drr_check_view(ctx, vols, view);
    if (rc) return rc;
    if (!intensity || !photon_prob) return drr_set_error(DRR_ERR_INVALID, "drr_project: NULL output");
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != ctx->device) DRR_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    ViewDev vd;
    view_to_dev(*view, vols, vd);
    Geo1 geo;
    view_to_geo1(*view, geo);
    const VolDev vol0 = vol_to_dev(vols[0]);
    SpectrumDev sp{ctx->d_energies, ctx->d_pdf, ctx->d_mu, ctx->n_bins};
    OutDev o{intensity, photon_prob, area_density, solid_angle_out};
    const drr_scene_t &sc = ctx->scene;
    const int tx = (view->width + 15) / 16, ty = (view->height + 15) / 16;
    dim3 grid((unsigned)(tx * ty)), block(256);
    (void)hipEventRecord(ctx->ev0, st);
    if (view->n_volumes == 1) {
        switch (sc.n_materials) {
#define DRR_CASE(MM) \
    case MM:                                                                                           \
        launch_v1<MM>(sc.attenuate_outside != 0, sc.march == DRR_MARCH_FAST, grid, block, st, geo, vol0, sp, \
                      sc.step, sc.air_index, o, tx);                                                     \
        break;
            DRR_CASE(1) DRR_CASE(2) DRR_CASE(3) DRR_CASE(4) DRR_CASE(5) DRR_CASE(6) DRR_CASE(7) DRR_CASE(8) \
        DRR_CASE(9) DRR_CASE(10) DRR_CASE(11) DRR_CASE(12) DRR_CASE(13) DRR_CASE(14) DRR_CASE(15) DRR_CASE(16)
#undef DRR_CASE
        }
    } else if (view->n_volumes == 2) {
        launch_multi<2>(sc.n_materials, grid, block, st, vd, sp, sc.step, sc.attenuate_outside, sc.air_index, o, tx);
    } else if (view->n_volumes == 3) {
        launch_multi<3>(sc.n_materials, grid, block, st, vd, sp, sc.step, sc.attenuate_outside, sc.air_index, o, tx);
    } else {
        launch_multi<4>(sc.n_materials, grid, block, st, vd, sp, sc.step, sc.attenuate_outside, sc.air_index, o, tx);
    }
    hipError_t e = hipGetLastError();
    (void)hipEventRecord(ctx->ev1, st);
    ctx->timed = true;
    if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
    if (e != hipSuccess) return drr_hip_error(e, "drr_project launch");
    return DRR_OK;
}
