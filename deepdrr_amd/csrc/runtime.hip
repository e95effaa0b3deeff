// runtime.hip -- context, spectrum tables, volume upload and error plumbing
// of the C ABI (include/deepdrr_amd.h).
//
// Replaces the reference's CuPy device setup and texture creation
// (deepdrr/projector/projector.py:116-257, 1395-1717, 1719-1764).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "drr_internal.h"

static thread_local std::string g_last_error;

int drr_set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int drr_hip_error(hipError_t e, const char *what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return DRR_ERR_HIP;
}

namespace {

// Keeps the caller's current device and switches to `dev` for the scope.
struct DeviceGuard {
    int prev = -1;
    bool ok = true;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) ok = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

// Device buffer freed at scope exit unless released: every early return of
// an entry point frees what it allocated so far.
struct DevBuf {
    void *p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes); }
};

}  // namespace

extern "C" {

const char *drr_last_error(void) { return g_last_error.c_str(); }

int drr_device_count(int *count) {
    if (!count) return drr_set_error(DRR_ERR_INVALID, "drr_device_count: count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return DRR_OK;
}

int drr_ctx_create(int device_id, drr_ctx **out) {
    if (!out) return drr_set_error(DRR_ERR_INVALID, "drr_ctx_create: out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return drr_set_error(DRR_ERR_UNSUPPORTED, "drr_ctx_create: no HIP device available");
    if (device_id < 0 || device_id >= n)
        return drr_set_error(DRR_ERR_INVALID, "drr_ctx_create: device id out of range");
    DeviceGuard g(device_id);
    if (!g.ok) return drr_set_error(DRR_ERR_HIP, "drr_ctx_create: hipSetDevice failed");
    drr_ctx *c = new drr_ctx();
    c->device = device_id;
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return drr_set_error(DRR_ERR_HIP, "drr_ctx_create: hipEventCreate failed");
    }
    *out = c;
    return DRR_OK;
}

int drr_ctx_destroy(drr_ctx *ctx) {
    if (!ctx) return DRR_OK;
    DeviceGuard g(ctx->device);
    (void)hipFree(ctx->d_energies);
    (void)hipFree(ctx->d_pdf);
    (void)hipFree(ctx->d_mu);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    delete ctx;
    return DRR_OK;
}

int drr_ctx_set_scene(drr_ctx *ctx, const drr_scene_t *scene) {
    if (!ctx || !scene) return drr_set_error(DRR_ERR_INVALID, "drr_ctx_set_scene: NULL argument");
    if (!(scene->step > 0.0f) || !std::isfinite(scene->step))
        return drr_set_error(DRR_ERR_INVALID, "drr_ctx_set_scene: step must be a positive finite number");
    if (scene->n_materials < 1 || scene->n_materials > DRR_MAX_MATERIALS)
        return drr_set_error(DRR_ERR_INVALID, "drr_ctx_set_scene: n_materials must be in [1, 16]");
    if (scene->air_index < 0 || scene->air_index >= scene->n_materials)
        return drr_set_error(DRR_ERR_INVALID, "drr_ctx_set_scene: air_index out of range");
    if (scene->march != DRR_MARCH_EXACT && scene->march != DRR_MARCH_FAST)
        return drr_set_error(DRR_ERR_INVALID, "drr_ctx_set_scene: march must be DRR_MARCH_EXACT or DRR_MARCH_FAST");
    if (ctx->have_scene && ctx->n_bins > 0 && scene->n_materials != ctx->scene.n_materials) {
        ctx->n_bins = 0;  // mu table no longer matches; caller must set the spectrum again
    }
    ctx->scene = *scene;
    ctx->have_scene = true;
    return DRR_OK;
}

int drr_ctx_set_spectrum(drr_ctx *ctx, int n_bins, const float *energies_keV, const float *pdf,
                         const float *mu_over_rho) {
    if (!ctx || !energies_keV || !pdf || !mu_over_rho)
        return drr_set_error(DRR_ERR_INVALID, "drr_ctx_set_spectrum: NULL argument");
    if (!ctx->have_scene)
        return drr_set_error(DRR_ERR_STATE, "drr_ctx_set_spectrum: call drr_ctx_set_scene first");
    if (n_bins < 1 || n_bins > DRR_MAX_BINS)
        return drr_set_error(DRR_ERR_INVALID, "drr_ctx_set_spectrum: n_bins must be in [1, 256]");
    DeviceGuard g(ctx->device);
    const int M = ctx->scene.n_materials;
    (void)hipFree(ctx->d_energies);
    (void)hipFree(ctx->d_pdf);
    (void)hipFree(ctx->d_mu);
    ctx->d_energies = ctx->d_pdf = ctx->d_mu = nullptr;
    ctx->n_bins = 0;
    DRR_HIP_CHECK(hipMalloc(&ctx->d_energies, sizeof(float) * n_bins));
    DRR_HIP_CHECK(hipMalloc(&ctx->d_pdf, sizeof(float) * n_bins));
    DRR_HIP_CHECK(hipMalloc(&ctx->d_mu, sizeof(float) * n_bins * M));
    DRR_HIP_CHECK(hipMemcpy(ctx->d_energies, energies_keV, sizeof(float) * n_bins, hipMemcpyHostToDevice));
    DRR_HIP_CHECK(hipMemcpy(ctx->d_pdf, pdf, sizeof(float) * n_bins, hipMemcpyHostToDevice));
    DRR_HIP_CHECK(hipMemcpy(ctx->d_mu, mu_over_rho, sizeof(float) * n_bins * M, hipMemcpyHostToDevice));
    ctx->n_bins = n_bins;
    return DRR_OK;
}

int drr_ctx_last_kernel_ms(drr_ctx *ctx, float *ms) {
    if (!ctx || !ms) return drr_set_error(DRR_ERR_INVALID, "drr_ctx_last_kernel_ms: NULL argument");
    if (!ctx->timed) return drr_set_error(DRR_ERR_STATE, "drr_ctx_last_kernel_ms: nothing timed yet");
    DeviceGuard g(ctx->device);
    DRR_HIP_CHECK(hipEventSynchronize(ctx->ev1));
    DRR_HIP_CHECK(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return DRR_OK;
}

// ---- volumes --------------------------------------------------------------

__global__ void k_label_max(const uint8_t *labels, int64_t n, unsigned int *out) {
    unsigned int m = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        m = max(m, (unsigned int)labels[i]);
    atomicMax(out, m);
}

__global__ void k_hu_ingest(const int16_t *hu, int64_t n, int n_classes, const float *upper,
                            const uint8_t *label_of_class, float *density, uint8_t *labels) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double h = (double)hu[i];
        // volume.py:348-350 in FP64, rounded once to float32 (volume.py:182)
        double a = 0.001029 * h + 1.030;
        double b = 0.0005886 * h + 1.03;
        double r = fmax(fmin(a, b), 0.0);
        density[i] = (float)r;
        uint8_t lab = label_of_class[n_classes - 1];
        for (int c = 0; c < n_classes; c++) {
            if ((float)hu[i] <= upper[c]) {
                lab = label_of_class[c];
                break;
            }
        }
        labels[i] = lab;
    }
}

static int volume_alloc(drr_ctx *ctx, int nx, int ny, int nz, drr_volume **out) {
    drr_volume *v = new drr_volume();
    v->ctx = ctx;
    v->device = ctx->device;
    v->nx = nx;
    v->ny = ny;
    v->nz = nz;
    v->n_cells = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    hipError_t e = hipMalloc(&v->cells, sizeof(uint32_t) * DRR_CELL_WORDS * v->n_cells);
    if (e != hipSuccess) {
        (void)hipFree(v->cells);
        delete v;
        return drr_set_error(DRR_ERR_ALLOC, std::string("drr_volume: device allocation failed: ") +
                                                hipGetErrorString(e));
    }
    *out = v;
    return DRR_OK;
}

static int check_dims(int nx, int ny, int nz) {
    if (nx < 1 || ny < 1 || nz < 1) return drr_set_error(DRR_ERR_INVALID, "drr_volume: dimensions must be >= 1");
    if (nx > 16384 || ny > 16384 || nz > 16384)
        return drr_set_error(DRR_ERR_INVALID, "drr_volume: dimensions must be <= 16384");
    // the march forms cell indices in 32 bits, biased by 0x4B000000 (below
    // 2^32 for fewer than 2^31 cells); byte offsets are 64-bit
    if ((int64_t)(nx + 1) * (ny + 1) * (nz + 1) >= ((int64_t)1 << 31))
        return drr_set_error(DRR_ERR_INVALID, "drr_volume: (nx+1)(ny+1)(nz+1) must be < 2^31 cells");
    return DRR_OK;
}

int drr_volume_create(drr_ctx *ctx, const float *density, const uint8_t *labels, int nx, int ny,
                      int nz, int on_device, drr_volume **out) {
    if (!ctx || !density || !labels || !out)
        return drr_set_error(DRR_ERR_INVALID, "drr_volume_create: NULL argument");
    *out = nullptr;
    if (!ctx->have_scene) return drr_set_error(DRR_ERR_STATE, "drr_volume_create: call drr_ctx_set_scene first");
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    DeviceGuard g(ctx->device);
    const int64_t n = (int64_t)nx * ny * nz;
    const int M = ctx->scene.n_materials;
    const float *d_den = density;
    const uint8_t *d_lab = labels;
    DevBuf tmp_den, tmp_lab, d_max;
    if (!on_device) {
        for (int64_t i = 0; i < n; i++)
            if (labels[i] >= M)
                return drr_set_error(DRR_ERR_INVALID, "drr_volume_create: label >= n_materials at voxel " +
                                                          std::to_string(i));
        DRR_HIP_CHECK(tmp_den.alloc(sizeof(float) * n));
        DRR_HIP_CHECK(tmp_lab.alloc(n));
        DRR_HIP_CHECK(hipMemcpy(tmp_den.p, density, sizeof(float) * n, hipMemcpyHostToDevice));
        DRR_HIP_CHECK(hipMemcpy(tmp_lab.p, labels, n, hipMemcpyHostToDevice));
        d_den = tmp_den.as<float>();
        d_lab = tmp_lab.as<uint8_t>();
    } else {
        // device inputs may come from any stream of the caller: wait for all
        // queued work before reading them
        DRR_HIP_CHECK(hipDeviceSynchronize());
        DRR_HIP_CHECK(d_max.alloc(sizeof(unsigned int)));
        DRR_HIP_CHECK(hipMemset(d_max.p, 0, sizeof(unsigned int)));
        k_label_max<<<1024, 256, 0, 0>>>(labels, n, d_max.as<unsigned int>());
        DRR_HIP_CHECK(hipGetLastError());
        unsigned int mx = 0;
        DRR_HIP_CHECK(hipMemcpy(&mx, d_max.p, sizeof(unsigned int), hipMemcpyDeviceToHost));
        if ((int)mx >= M)
            return drr_set_error(DRR_ERR_INVALID, "drr_volume_create: label >= n_materials");
    }
    drr_volume *v = nullptr;
    rc = volume_alloc(ctx, nx, ny, nz, &v);
    if (rc == DRR_OK) rc = drr_launch_expand(d_den, d_lab, nx, ny, nz, v->cells, 0);
    if (rc == DRR_OK) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = drr_hip_error(e, "drr_volume_create: expand");
    }
    if (rc != DRR_OK) {
        if (v) drr_volume_destroy(v);
        return rc;
    }
    *out = v;
    return DRR_OK;
}

int drr_volume_create_from_hu(drr_ctx *ctx, const int16_t *hu, int nx, int ny, int nz, int on_device,
                              int n_classes, const float *upper_hu, const uint8_t *label_of_class,
                              drr_volume **out) {
    if (!ctx || !hu || !upper_hu || !label_of_class || !out)
        return drr_set_error(DRR_ERR_INVALID, "drr_volume_create_from_hu: NULL argument");
    *out = nullptr;
    if (!ctx->have_scene)
        return drr_set_error(DRR_ERR_STATE, "drr_volume_create_from_hu: call drr_ctx_set_scene first");
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    if (n_classes < 1 || n_classes > 32)
        return drr_set_error(DRR_ERR_INVALID, "drr_volume_create_from_hu: n_classes must be in [1, 32]");
    for (int c = 0; c < n_classes; c++)
        if (label_of_class[c] >= ctx->scene.n_materials)
            return drr_set_error(DRR_ERR_INVALID, "drr_volume_create_from_hu: label_of_class >= n_materials");
    DeviceGuard g(ctx->device);
    const int64_t n = (int64_t)nx * ny * nz;
    DevBuf tmp_hu, d_den, d_lab, d_upper, d_map;
    const int16_t *src = hu;
    if (!on_device) {
        DRR_HIP_CHECK(tmp_hu.alloc(sizeof(int16_t) * n));
        DRR_HIP_CHECK(hipMemcpy(tmp_hu.p, hu, sizeof(int16_t) * n, hipMemcpyHostToDevice));
        src = tmp_hu.as<int16_t>();
    } else {
        DRR_HIP_CHECK(hipDeviceSynchronize());  // inputs from any stream of the caller
    }
    DRR_HIP_CHECK(d_den.alloc(sizeof(float) * n));
    DRR_HIP_CHECK(d_lab.alloc(n));
    DRR_HIP_CHECK(d_upper.alloc(sizeof(float) * n_classes));
    DRR_HIP_CHECK(d_map.alloc(n_classes));
    DRR_HIP_CHECK(hipMemcpy(d_upper.p, upper_hu, sizeof(float) * n_classes, hipMemcpyHostToDevice));
    DRR_HIP_CHECK(hipMemcpy(d_map.p, label_of_class, n_classes, hipMemcpyHostToDevice));
    k_hu_ingest<<<2048, 256, 0, 0>>>(src, n, n_classes, d_upper.as<float>(), d_map.as<uint8_t>(),
                                     d_den.as<float>(), d_lab.as<uint8_t>());
    DRR_HIP_CHECK(hipGetLastError());
    drr_volume *v = nullptr;
    rc = volume_alloc(ctx, nx, ny, nz, &v);
    if (rc == DRR_OK) rc = drr_launch_expand(d_den.as<float>(), d_lab.as<uint8_t>(), nx, ny, nz, v->cells, 0);
    if (rc == DRR_OK) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = drr_hip_error(e, "drr_volume_create_from_hu");
    }
    if (rc != DRR_OK) {
        if (v) drr_volume_destroy(v);
        return rc;
    }
    *out = v;
    return DRR_OK;
}

int drr_volume_destroy(drr_volume *vol) {
    if (!vol) return DRR_OK;
    DeviceGuard g(vol->device);
    (void)hipFree(vol->cells);
    (void)hipFree(vol->lin_vox);
    delete vol;
    return DRR_OK;
}

__global__ void k_unexpand(const uint32_t *cells, int nx, int ny, int nz, float *out_d, uint8_t *out_l) {
    const int64_t n = (int64_t)nx * ny * nz;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx);
        const int y = (int)((i / nx) % ny);
        const int z = (int)(i / ((int64_t)nx * ny));
        const int64_t cell = ((int64_t)(z + 1) * (ny + 1) + (y + 1)) * (int64_t)(nx + 1) + (x + 1);
        const uint32_t *rec = cells + DRR_CELL_WORDS * cell;  // corner 0 of the voxel's own cell
        out_d[i] = __uint_as_float(rec[0]);
        out_l[i] = (uint8_t)(rec[8] & 0xfu);
    }
}

}  // extern "C"

// Voxel-order (density bits, label) records from the cells, and the largest
// density (each thread's maximum starts at 0, so it is non-negative, and
// non-negative floats order like their bits).
__global__ void k_voxels(const uint32_t *cells, int nx, int ny, int nz, uint2 *out, unsigned int *dmax) {
    const int64_t n = (int64_t)nx * ny * nz;
    float m = 0.0f;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx);
        const int y = (int)((i / nx) % ny);
        const int z = (int)(i / ((int64_t)nx * ny));
        const int64_t cell = ((int64_t)(z + 1) * (ny + 1) + (y + 1)) * (int64_t)(nx + 1) + (x + 1);
        const uint32_t *rec = cells + DRR_CELL_WORDS * cell;  // corner 0 of the voxel's own cell
        out[i] = make_uint2(rec[0], rec[8] & 0xfu);
        m = fmaxf(m, __uint_as_float(rec[0]));
    }
    atomicMax(dmax, __float_as_uint(m));
}

int drr_volume_ensure_linear(drr_volume *vol) {
    if (vol->lin_vox) return DRR_OK;
    const int64_t n = (int64_t)vol->nx * vol->ny * vol->nz;
    uint2 *vox = nullptr;
    unsigned int *d_max = nullptr;
    hipError_t e = hipMalloc(&vox, sizeof(uint2) * n);
    if (e == hipSuccess) e = hipMalloc(&d_max, sizeof(unsigned int));
    if (e == hipSuccess) e = hipMemset(d_max, 0, sizeof(unsigned int));
    if (e == hipSuccess) {
        k_voxels<<<2048, 256, 0, 0>>>(vol->cells, vol->nx, vol->ny, vol->nz, vox, d_max);
        e = hipGetLastError();
    }
    unsigned int bits = 0;
    if (e == hipSuccess) e = hipMemcpy(&bits, d_max, sizeof(unsigned int), hipMemcpyDeviceToHost);
    (void)hipFree(d_max);
    if (e != hipSuccess) {
        (void)hipFree(vox);
        return drr_hip_error(e, "voxel-order copy for the Monte-Carlo path");
    }
    float mx;
    std::memcpy(&mx, &bits, sizeof(float));
    vol->lin_vox = vox;
    vol->max_density = mx;
    return DRR_OK;
}

extern "C" {

int drr_volume_download(drr_volume *vol, float *density, uint8_t *labels) {
    if (!vol || !density || !labels) return drr_set_error(DRR_ERR_INVALID, "drr_volume_download: NULL argument");
    DeviceGuard g(vol->device);
    const int64_t n = (int64_t)vol->nx * vol->ny * vol->nz;
    DevBuf d, l;
    DRR_HIP_CHECK(d.alloc(sizeof(float) * n));
    DRR_HIP_CHECK(l.alloc(n));
    k_unexpand<<<2048, 256, 0, 0>>>(vol->cells, vol->nx, vol->ny, vol->nz, d.as<float>(), l.as<uint8_t>());
    DRR_HIP_CHECK(hipGetLastError());
    DRR_HIP_CHECK(hipMemcpy(density, d.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    DRR_HIP_CHECK(hipMemcpy(labels, l.p, n, hipMemcpyDeviceToHost));
    return DRR_OK;
}

}  // extern "C"

int drr_check_view(const drr_ctx *ctx, const drr_volume *const *vols, const drr_view_t *view) {
    if (!ctx || !view) return drr_set_error(DRR_ERR_INVALID, "NULL context or view");
    if (!ctx->have_scene) return drr_set_error(DRR_ERR_STATE, "scene not set (drr_ctx_set_scene)");
    if (ctx->n_bins < 1) return drr_set_error(DRR_ERR_STATE, "spectrum not set (drr_ctx_set_spectrum)");
    if (view->n_volumes < 1 || view->n_volumes > DRR_MAX_VOLUMES)
        return drr_set_error(DRR_ERR_INVALID, "view.n_volumes must be in [1, 4]");
    if (view->width < 1 || view->height < 1 || view->width > 65536 || view->height > 65536)
        return drr_set_error(DRR_ERR_INVALID, "view width/height must be in [1, 65536]");
    for (int i = 0; i < view->n_volumes; i++) {
        if (!vols || !vols[i]) return drr_set_error(DRR_ERR_INVALID, "NULL volume in view");
        if (vols[i]->ctx != ctx) return drr_set_error(DRR_ERR_INVALID, "volume belongs to another context");
    }
    return drr_check_view_finite(view);
}

int drr_check_view_finite(const drr_view_t *view) {
    // the march's cell indices are clamped from float sample positions, which
    // stay finite (and in range after the clamp) only for finite geometry
    bool ok = std::isfinite(view->max_ray_length);
    for (int i = 0; i < 9; i++) ok = ok && std::isfinite(view->world_from_index[i]);
    for (int v = 0; v < view->n_volumes && v < DRR_MAX_VOLUMES; v++) {
        const drr_view_volume_t &q = view->vol[v];
        for (int i = 0; i < 12; i++) ok = ok && std::isfinite(q.ijk_from_world[i]);
        for (int i = 0; i < 3; i++)
            ok = ok && std::isfinite(q.source_ijk[i]) && std::isfinite(q.bounds_lo[i]) && std::isfinite(q.bounds_hi[i]);
    }
    return ok ? DRR_OK : drr_set_error(DRR_ERR_INVALID, "view geometry must be finite");
}
