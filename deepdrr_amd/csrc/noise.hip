// noise.hip -- Poisson shot noise with the correlated 3x3 kernel, on the device
// (deepdrr/projector/analytic_generators.py:10-18):
//   lambda = photon_prob * photon_count
//   shot   = (Poisson(lambda) - lambda) * intensity / lambda
//   out    = clip(intensity + conv2d_same(shot, K) + 0 (electronic), 0, 1e31)
// K = [[0.03, 0.06, 0.02], [0.11, 0.98, 0.11], [0.02, 0.06, 0.03]] (Wang et al.
// 2014, as in the reference); K is centrosymmetric, so the convolution equals
// the correlation computed here, with zero padding ('same').
// Random numbers: Philox4x32-10 (Salmon et al. 2011) keyed by the seed, with
// the counter (pixel, image, draw); Poisson variates by inversion for
// lambda < 10 and Hoermann's PTRS transformed rejection (1993) above.  The
// reference draws from NumPy's global generator; results agree in
// distribution, not value.
#include <hip/hip_runtime.h>

#include "drr_internal.h"

namespace {

struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox(U4 c, uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

struct Stream {
    uint32_t pix, img, draw;
    uint32_t k0, k1;
    uint32_t buf[4];
    int left = 0;
    __device__ float uniform() {  // (0, 1)
        if (left == 0) {
            U4 r = philox(U4{pix, img, draw++, 0x5eedu}, k0, k1);
            buf[0] = r.x;
            buf[1] = r.y;
            buf[2] = r.z;
            buf[3] = r.w;
            left = 4;
        }
        const uint32_t v = buf[--left];
        return ((float)(v >> 8) + 0.5f) * (1.0f / 16777216.0f);
    }
};

__device__ float poisson(Stream &s, float lam) {
    if (!(lam > 0.0f)) return 0.0f;
    if (lam < 10.0f) {
        const float L = expf(-lam);
        float p = 1.0f;
        int k = 0;
        for (int it = 0; it < 1000; it++) {
            p *= s.uniform();
            if (p <= L) break;
            k++;
        }
        return (float)k;
    }
    const float slam = sqrtf(lam);
    const double loglam_d = log((double)lam);
    const float b = 0.931f + 2.53f * slam;
    const float a = -0.059f + 0.02483f * b;
    const float invalpha = 1.1239f + 1.1328f / (b - 3.4f);
    const float vr = 0.9277f - 3.6224f / (b - 2.0f);
    for (int it = 0; it < 1000; it++) {
        const float U = s.uniform() - 0.5f;
        const float V = s.uniform();
        const float us = 0.5f - fabsf(U);
        const float k = floorf((2.0f * a / us + b) * U + lam + 0.43f);
        if (us >= 0.07f && V <= vr) return k;
        if (k < 0.0f || (us < 0.013f && V > us)) continue;
        // log of the Poisson pmf at k: its three terms reach ~1e7 at lam ~ 1e6
        // and cancel, so they are summed in FP64
        const double logpmf = -(double)lam + (double)k * (double)loglam_d - lgamma((double)k + 1.0);
        if ((double)(logf(V) + logf(invalpha) - logf(a / (us * us) + b)) <= logpmf) return k;
    }
    return roundf(lam);
}

}  // namespace

__global__ __launch_bounds__(256) void k_shot(const float *I, const float *P, float *shot, int n_images,
                                              int64_t npix, float count, uint32_t k0, uint32_t k1) {
    const int64_t total = (int64_t)n_images * npix;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float lam = P[i] * count;
        Stream s;
        s.pix = (uint32_t)(i % npix);
        s.img = (uint32_t)(i / npix);
        s.draw = 0;
        s.k0 = k0;
        s.k1 = k1;
        const float k = poisson(s, lam);
        shot[i] = (lam > 0.0f) ? (k - lam) * I[i] / lam : 0.0f;
    }
}

__global__ __launch_bounds__(256) void k_correlate(const float *I, const float *shot, float *out,
                                                   int n_images, int H, int W) {
    const float K[3][3] = {{0.03f, 0.06f, 0.02f}, {0.11f, 0.98f, 0.11f}, {0.02f, 0.06f, 0.03f}};
    const int64_t npix = (int64_t)H * W, total = (int64_t)n_images * npix;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t base = (i / npix) * npix;
        const int y = (int)((i % npix) / W), x = (int)(i % W);
        float acc = 0.0f;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const int yy = y + 1 - a, xx = x + 1 - b;  // convolution: flipped offsets
                if (yy >= 0 && yy < H && xx >= 0 && xx < W) acc += shot[base + (int64_t)yy * W + xx] * K[a][b];
            }
        out[i] = fminf(fmaxf(I[i] + acc, 0.0f), 1e31f);
    }
}

extern "C" int drr_add_noise(drr_ctx *ctx, const float *intensity, const float *photon_prob, float *out,
                             int n_images, int height, int width, float photon_count, uint64_t seed,
                             void *stream) {
    if (!ctx || !intensity || !photon_prob || !out)
        return drr_set_error(DRR_ERR_INVALID, "drr_add_noise: NULL argument");
    if (n_images < 1 || height < 1 || width < 1 || !(photon_count > 0.0f))
        return drr_set_error(DRR_ERR_INVALID, "drr_add_noise: bad sizes or photon_count");
    if (out == intensity) return drr_set_error(DRR_ERR_INVALID, "drr_add_noise: out must not alias intensity");
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != ctx->device) DRR_HIP_CHECK(hipSetDevice(ctx->device));
    const int64_t total = (int64_t)n_images * height * width;
    float *shot = nullptr;
    DRR_HIP_CHECK(hipMalloc(&shot, sizeof(float) * total));
    hipStream_t st = (hipStream_t)stream;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    k_shot<<<(unsigned)blocks, 256, 0, st>>>(intensity, photon_prob, shot, n_images, (int64_t)height * width,
                                             photon_count, (uint32_t)seed, (uint32_t)(seed >> 32));
    k_correlate<<<(unsigned)blocks, 256, 0, st>>>(intensity, shot, out, n_images, height, width);
    hipError_t e = hipGetLastError();
    hipError_t e2 = hipStreamSynchronize(st);
    (void)hipFree(shot);
    if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
    if (e != hipSuccess) return drr_hip_error(e, "drr_add_noise launch");
    if (e2 != hipSuccess) return drr_hip_error(e2, "drr_add_noise");
    return DRR_OK;
}
