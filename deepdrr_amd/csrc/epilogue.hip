// epilogue.hip -- device image epilogue: per-image neglog
// (deepdrr/utils/image_utils.py:18-59) over a set of images at arbitrary
// offsets, as used by the batched path and drr_neglog.
#include <hip/hip_runtime.h>

#include <vector>

#include "drr_internal.h"
#include "epilogue.h"

// Order-preserving float <-> uint mapping for atomic min/max.
__device__ __forceinline__ unsigned int f2key(float f) {
    const unsigned int b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key2f(unsigned int k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// keys[4*i + 0] = min x, [1] = min y, [2] = max y  (image i)
__global__ void k_neglog_reset(unsigned int *keys, int n_images) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_images; i += gridDim.x * blockDim.x) {
        keys[4 * i + 0] = 0xffffffffu;
        keys[4 * i + 1] = 0xffffffffu;
        keys[4 * i + 2] = 0u;
        keys[4 * i + 3] = 0u;
    }
}

// grid.y = image, grid.x = chunks of that image
__global__ __launch_bounds__(256) void k_img_min(const float *in, const ImgSeg *segs, unsigned int *keys) {
    const ImgSeg s = segs[blockIdx.y];
    float m = INFINITY;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < s.npix;
         i += (int64_t)gridDim.x * blockDim.x)
        m = fminf(m, in[s.offset + i]);
    m = wave_min(m);
    if ((threadIdx.x & 63) == 0 && m < INFINITY) atomicMin(&keys[4 * blockIdx.y], f2key(m));
}

__global__ __launch_bounds__(256) void k_neglog_y(const float *in, float *out, const ImgSeg *segs,
                                                  unsigned int *keys, float eps) {
    const ImgSeg s = segs[blockIdx.y];
    // image += image.min() + epsilon  (float32: t = fl(min + fl(eps)))
    const float t = key2f(keys[4 * blockIdx.y]) + eps;
    float lo = INFINITY, hi = -INFINITY;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < s.npix;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float y = -logf(in[s.offset + i] + t);
        out[s.offset + i] = y;
        lo = fminf(lo, y);
        hi = fmaxf(hi, y);
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    if ((threadIdx.x & 63) == 0) {
        if (lo < INFINITY) atomicMin(&keys[4 * blockIdx.y + 1], f2key(lo));
        if (hi > -INFINITY) atomicMax(&keys[4 * blockIdx.y + 2], f2key(hi));
    }
}

// keys[3] = 1 if any image is constant (min y == max y).
__global__ __launch_bounds__(256) void k_neglog_flag(unsigned int *keys, int n_images) {
    unsigned int any = 0;
    for (int j = threadIdx.x; j < n_images; j += blockDim.x) any |= (keys[4 * j + 1] == keys[4 * j + 2]);
    __shared__ unsigned int sh;
    if (threadIdx.x == 0) sh = 0;
    __syncthreads();
    if (any) atomicOr(&sh, 1u);
    __syncthreads();
    if (threadIdx.x == 0) keys[3] = sh;
}

// (y - min) / (max - min); if any image is constant every image becomes 0,
// as the reference does (image_utils.py:43-51).
__global__ __launch_bounds__(256) void k_neglog_norm(float *out, const ImgSeg *segs,
                                                     const unsigned int *keys) {
    const bool any_const = keys[3] != 0;
    const ImgSeg s = segs[blockIdx.y];
    const float lo = key2f(keys[4 * blockIdx.y + 1]);
    const float hi = key2f(keys[4 * blockIdx.y + 2]);
    const float den = hi - lo;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < s.npix;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float y = out[s.offset + i];
        out[s.offset + i] = any_const ? 0.0f : (y - lo) / den;
    }
}

int drr_enqueue_neglog(const float *in, float *out, const ImgSeg *d_segs, int n_images,
                       int64_t max_npix, unsigned int *d_keys, float eps, hipStream_t st) {
    int chunks = (int)((max_npix + 256 * 16 - 1) / (256 * 16));
    if (chunks < 1) chunks = 1;
    if (chunks > 1024) chunks = 1024;
    dim3 grid(chunks, n_images);
    k_neglog_reset<<<(n_images + 255) / 256, 256, 0, st>>>(d_keys, n_images);
    k_img_min<<<grid, 256, 0, st>>>(in, d_segs, d_keys);
    k_neglog_y<<<grid, 256, 0, st>>>(in, out, d_segs, d_keys, eps);
    k_neglog_flag<<<1, 256, 0, st>>>(d_keys, n_images);
    k_neglog_norm<<<grid, 256, 0, st>>>(out, d_segs, d_keys);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return drr_hip_error(e, "neglog launch");
    return DRR_OK;
}

extern "C" int drr_neglog(drr_ctx *ctx, const float *in, float *out, int n_images, int64_t npix,
                          float epsilon, void *stream) {
    if (!ctx || !in || !out) return drr_set_error(DRR_ERR_INVALID, "drr_neglog: NULL argument");
    if (n_images < 1 || npix < 1 || n_images > 65535)
        return drr_set_error(DRR_ERR_INVALID, "drr_neglog: need 1..65535 images of >= 1 pixel");
    hipStream_t st = (hipStream_t)stream;
    ImgSeg *segs = nullptr;
    unsigned int *keys = nullptr;
    std::vector<ImgSeg> h(n_images);
    for (int i = 0; i < n_images; i++) h[i] = ImgSeg{(int64_t)i * npix, npix};
    int rc = DRR_OK;
    hipError_t e = hipMalloc(&segs, sizeof(ImgSeg) * n_images);
    if (e == hipSuccess) e = hipMalloc(&keys, sizeof(unsigned int) * 4 * n_images);
    if (e == hipSuccess) e = hipMemcpyAsync(segs, h.data(), sizeof(ImgSeg) * n_images, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) rc = drr_hip_error(e, "drr_neglog setup");
    if (rc == DRR_OK) rc = drr_enqueue_neglog(in, out, segs, n_images, npix, keys, epsilon, st);
    e = hipStreamSynchronize(st);
    if (rc == DRR_OK && e != hipSuccess) rc = drr_hip_error(e, "drr_neglog");
    (void)hipFree(segs);
    (void)hipFree(keys);
    return rc;
}
