// batch.hip -- batched forward projection: hundreds of (volume, pose,
// detector) jobs in one persistent launch, replayed from a hipGraph.
//
// This is the added batched form of deepdrr's per-view loop
// (deepdrr/projector/projector.py:679-685 launching projectKernel once per
// view).  Each job is one view of one volume; its detector is cut into ray
// tiles of 64 pixels (one wavefront each).  The tile list is grouped by
// volume and job and split into 8 contiguous segments, one per XCD: a
// persistent wave pulls tiles from its own XCD's segment (atomic counter) and
// steals from the others once it is empty, so each XCD's L2 works on a
// coherent range of poses/rays and the tail is balanced at tile granularity.
// The per-ray march is the same device code as drr_project, so every image is
// bit-identical to the one-view path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "drr_internal.h"
#include "epilogue.h"
#include "march.h"

using namespace drr;

struct JobDev {
    Geo1 g;
    int vol;
    int tiles_x;
    int64_t out_offset;
};

struct drr_batch_s {
    drr_ctx *ctx = nullptr;
    int n_jobs = 0, n_vols = 0;
    int M = 0;
    bool atten = false;
    drr_batch_tuning_t tune{};
    JobDev *d_jobs = nullptr;
    VolDev *d_vols = nullptr;
    int2 *d_items = nullptr;
    int *d_seg = nullptr;        // 9 segment starts
    int *d_counters = nullptr;   // 8 queue counters
    ImgSeg *d_imgs = nullptr;
    unsigned int *d_keys = nullptr;
    int64_t n_items = 0, max_npix = 0;
    int grid = 0, block = 0;
    hipStream_t cap = nullptr;
    hipGraph_t graph_march = nullptr, graph_epi = nullptr;
    hipGraphExec_t exec_march = nullptr, exec_epi = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;  // before march, after march, after epilogue
    bool launched = false;
};

__global__ void k_batch_reset(int *counters) {
    if (threadIdx.x < 8) counters[threadIdx.x] = 0;
}

template <int M, bool ATTEN, int TW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MarchOcc<M + (ATTEN ? 1 : 0)>::waves))) void k_batch_march(const JobDev *__restrict__ jobs,
                                                     const VolDev *__restrict__ vols,
                                                     const int2 *__restrict__ items,
                                                     const int *__restrict__ seg, int *counters,
                                                     SpectrumDev sp, float step, int air,
                                                     float *__restrict__ I, float *__restrict__ P) {
    constexpr int TH = 64 / TW;
    const int lane = threadIdx.x & 63;
    int q = blockIdx.x & 7;  // blocks are dispatched round-robin over the 8 XCDs
    int exhausted = 0;
    while (exhausted < 8) {
        int pos = 0;
        if (lane == 0) pos = atomicAdd(&counters[q], 1);
        pos = __builtin_amdgcn_readfirstlane(__shfl(pos, 0));
        const int idx = __builtin_amdgcn_readfirstlane(seg[q] + pos);
        if (idx >= seg[q + 1]) {
            q = (q + 1) & 7;
            ++exhausted;
            continue;
        }
        const int2 it = items[idx];
        const int j = __builtin_amdgcn_readfirstlane(it.x);
        const int tile = __builtin_amdgcn_readfirstlane(it.y);
        const JobDev &J = jobs[j];
        const int tx = tile % J.tiles_x, ty = tile / J.tiles_x;
        const int udx = tx * TW + (lane % TW), vdx = ty * TH + (lane / TW);
        if (udx < J.g.width && vdx < J.g.height) {
            float ad[M];
            march_v1<M, ATTEN>(J.g, vols[J.vol], udx, vdx, step, air, ad);
            float inten, pp;
            spectral<M>(sp, ad, inten, pp);
            const int64_t o = J.out_offset + (int64_t)vdx * J.g.width + udx;
            I[o] = inten;
            P[o] = pp;
        }
    }
}

template <int M, bool ATTEN>
static void launch_march_tw(int tw, dim3 g, dim3 b, hipStream_t st, drr_batch *B, SpectrumDev sp,
                            float step, int air, float *I, float *P) {
    if (tw == 16)
        k_batch_march<M, ATTEN, 16><<<g, b, 0, st>>>(B->d_jobs, B->d_vols, B->d_items, B->d_seg,
                                                      B->d_counters, sp, step, air, I, P);
    else
        k_batch_march<M, ATTEN, 8><<<g, b, 0, st>>>(B->d_jobs, B->d_vols, B->d_items, B->d_seg,
                                                     B->d_counters, sp, step, air, I, P);
}

static void launch_march(drr_batch *B, hipStream_t st, SpectrumDev sp, float step, int air, float *I,
                         float *P) {
    dim3 g(B->grid), b(B->block);
    switch (B->M) {
#define DRR_CASE(MM)                                                                          \
    case MM:                                                                                  \
        if (B->atten)                                                                         \
            launch_march_tw<MM, true>(B->tune.tile_w, g, b, st, B, sp, step, air, I, P);     \
        else                                                                                  \
            launch_march_tw<MM, false>(B->tune.tile_w, g, b, st, B, sp, step, air, I, P);    \
        break;
        DRR_CASE(1) DRR_CASE(2) DRR_CASE(3) DRR_CASE(4) DRR_CASE(5) DRR_CASE(6) DRR_CASE(7) DRR_CASE(8)
#undef DRR_CASE
    }
}

static void batch_free(drr_batch *B) {
    if (!B) return;
    if (B->exec_march) (void)hipGraphExecDestroy(B->exec_march);
    if (B->exec_epi) (void)hipGraphExecDestroy(B->exec_epi);
    if (B->graph_march) (void)hipGraphDestroy(B->graph_march);
    if (B->graph_epi) (void)hipGraphDestroy(B->graph_epi);
    if (B->cap) (void)hipStreamDestroy(B->cap);
    for (hipEvent_t e : {B->ev0, B->ev1, B->ev2})
        if (e) (void)hipEventDestroy(e);
    (void)hipFree(B->d_jobs);
    (void)hipFree(B->d_vols);
    (void)hipFree(B->d_items);
    (void)hipFree(B->d_seg);
    (void)hipFree(B->d_counters);
    (void)hipFree(B->d_imgs);
    (void)hipFree(B->d_keys);
    delete B;
}

// Capture [queue reset, persistent march] into one graph and the neglog
// epilogue into a second, so ordinary events between the two replays time the
// march.  Captured work allocates nothing and records no memset nodes.
static hipError_t capture(drr_batch *B, SpectrumDev sp, float step, int air, float *I, float *P,
                          float *images, int neglog) {
    hipError_t e = hipStreamBeginCapture(B->cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    k_batch_reset<<<1, 64, 0, B->cap>>>(B->d_counters);
    launch_march(B, B->cap, sp, step, air, I, P);
    e = hipStreamEndCapture(B->cap, &B->graph_march);
    if (e != hipSuccess) return e;
    e = hipGraphInstantiate(&B->exec_march, B->graph_march, nullptr, nullptr, 0);
    if (e != hipSuccess || !neglog) return e;
    e = hipStreamBeginCapture(B->cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    int rc = drr_enqueue_neglog(I, images, B->d_imgs, B->n_jobs, B->max_npix, B->d_keys, 0.01f, B->cap);
    hipError_t e2 = hipStreamEndCapture(B->cap, &B->graph_epi);
    if (rc != DRR_OK) return hipErrorLaunchFailure;
    if (e2 != hipSuccess) return e2;
    return hipGraphInstantiate(&B->exec_epi, B->graph_epi, nullptr, nullptr, 0);
}

extern "C" int drr_batch_create(drr_ctx *ctx, const drr_volume *const *vols, int n_vols,
                                const drr_job_t *jobs, int n_jobs, const drr_batch_tuning_t *tuning,
                                float *intensity, float *photon_prob, float *images, int neglog,
                                drr_batch **out) {
    if (!ctx || !vols || !jobs || !out || !intensity || !photon_prob)
        return drr_set_error(DRR_ERR_INVALID, "drr_batch_create: NULL argument");
    *out = nullptr;
    if (neglog && !images) return drr_set_error(DRR_ERR_INVALID, "drr_batch_create: neglog needs images");
    if (n_vols < 1 || n_jobs < 1 || n_jobs > 65535)
        return drr_set_error(DRR_ERR_INVALID, "drr_batch_create: need >= 1 volume and 1..65535 jobs");
    if (!ctx->have_scene || ctx->n_bins < 1)
        return drr_set_error(DRR_ERR_STATE, "drr_batch_create: scene/spectrum not set");
    drr_batch_tuning_t tune{8, 8, 4, 8};
    if (tuning) tune = *tuning;
    if (!((tune.tile_w == 8 && tune.tile_h == 8) || (tune.tile_w == 16 && tune.tile_h == 4)))
        return drr_set_error(DRR_ERR_INVALID, "drr_batch_create: tile must be 8x8 or 16x4");
    if (tune.waves_per_block < 1 || tune.waves_per_block > 4 || tune.blocks_per_cu < 1 || tune.blocks_per_cu > 32)
        return drr_set_error(DRR_ERR_INVALID, "drr_batch_create: waves_per_block in [1,4], blocks_per_cu in [1,32]");
    for (int v = 0; v < n_vols; v++) {
        if (!vols[v] || vols[v]->ctx != ctx)
            return drr_set_error(DRR_ERR_INVALID, "drr_batch_create: NULL or foreign volume");
    }
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != ctx->device) DRR_HIP_CHECK(hipSetDevice(ctx->device));

    drr_batch *B = new drr_batch();
    B->ctx = ctx;
    B->n_jobs = n_jobs;
    B->n_vols = n_vols;
    B->M = ctx->scene.n_materials;
    B->atten = ctx->scene.attenuate_outside != 0;
    B->tune = tune;
    int rc = DRR_OK;
    auto fail = [&](int code) {
        batch_free(B);
        if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
        return code;
    };

    // Jobs and tiles.
    std::vector<JobDev> hj(n_jobs);
    std::vector<ImgSeg> himg(n_jobs);
    std::vector<int> order(n_jobs);
    for (int j = 0; j < n_jobs; j++) {
        const drr_job_t &J = jobs[j];
        if (J.view.n_volumes != 1)
            return fail(drr_set_error(DRR_ERR_UNSUPPORTED, "drr_batch_create: each job must view exactly one volume"));
        if (J.volume_index < 0 || J.volume_index >= n_vols)
            return fail(drr_set_error(DRR_ERR_INVALID, "drr_batch_create: job volume_index out of range"));
        if (J.view.width < 1 || J.view.height < 1 || J.out_offset < 0)
            return fail(drr_set_error(DRR_ERR_INVALID, "drr_batch_create: bad job size/offset"));
        view_to_geo1(J.view, hj[j].g);
        hj[j].vol = J.volume_index;
        hj[j].tiles_x = (J.view.width + tune.tile_w - 1) / tune.tile_w;
        hj[j].out_offset = J.out_offset;
        himg[j] = ImgSeg{J.out_offset, (int64_t)J.view.width * J.view.height};
        B->max_npix = std::max<int64_t>(B->max_npix, himg[j].npix);
        order[j] = j;
    }
    // Group tiles by volume (L2/MALL reuse across poses), then by job.
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return hj[a].vol < hj[b].vol; });
    std::vector<int2> items;
    for (int j : order) {
        const int ty = (jobs[j].view.height + tune.tile_h - 1) / tune.tile_h;
        const int n = hj[j].tiles_x * ty;
        for (int t = 0; t < n; t++) items.push_back(make_int2(j, t));
    }
    B->n_items = (int64_t)items.size();
    if (B->n_items >= (int64_t)1 << 31)
        return fail(drr_set_error(DRR_ERR_INVALID, "drr_batch_create: too many ray tiles"));
    int seg[9];
    for (int q = 0; q <= 8; q++) seg[q] = (int)(B->n_items * q / 8);
    std::vector<VolDev> hv(n_vols);
    for (int v = 0; v < n_vols; v++) hv[v] = vol_to_dev(vols[v]);

    hipError_t e = hipSuccess;
#define BCHK(x)                                                        \
    do {                                                               \
        e = (x);                                                       \
        if (e != hipSuccess) return fail(drr_hip_error(e, #x));        \
    } while (0)
    BCHK(hipMalloc(&B->d_jobs, sizeof(JobDev) * n_jobs));
    BCHK(hipMalloc(&B->d_vols, sizeof(VolDev) * n_vols));
    BCHK(hipMalloc(&B->d_items, sizeof(int2) * items.size()));
    BCHK(hipMalloc(&B->d_seg, sizeof(int) * 9));
    BCHK(hipMalloc(&B->d_counters, sizeof(int) * 8));
    BCHK(hipMalloc(&B->d_imgs, sizeof(ImgSeg) * n_jobs));
    BCHK(hipMalloc(&B->d_keys, sizeof(unsigned int) * 4 * n_jobs));
    BCHK(hipMemcpy(B->d_jobs, hj.data(), sizeof(JobDev) * n_jobs, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(B->d_vols, hv.data(), sizeof(VolDev) * n_vols, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(B->d_items, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice));
    BCHK(hipMemcpy(B->d_seg, seg, sizeof(int) * 9, hipMemcpyHostToDevice));
    BCHK(hipMemcpy(B->d_imgs, himg.data(), sizeof(ImgSeg) * n_jobs, hipMemcpyHostToDevice));
    int cus = 0;
    BCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    B->block = 64 * tune.waves_per_block;
    B->grid = cus * tune.blocks_per_cu;
    BCHK(hipStreamCreateWithFlags(&B->cap, hipStreamNonBlocking));
    BCHK(hipEventCreate(&B->ev0));
    BCHK(hipEventCreate(&B->ev1));
    BCHK(hipEventCreate(&B->ev2));
#undef BCHK
    SpectrumDev sp{ctx->d_energies, ctx->d_pdf, ctx->d_mu, ctx->n_bins};
    const float step = ctx->scene.step;
    const int air = ctx->scene.air_index;
    e = capture(B, sp, step, air, intensity, photon_prob, images, neglog);
    if (e != hipSuccess) return fail(drr_hip_error(e, "drr_batch_create: graph capture"));
    if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
    *out = B;
    return rc;
}

extern "C" int drr_batch_launch(drr_batch *B, void *stream) {
    if (!B) return drr_set_error(DRR_ERR_INVALID, "drr_batch_launch: NULL batch");
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != B->ctx->device) DRR_HIP_CHECK(hipSetDevice(B->ctx->device));
    hipStream_t st = (hipStream_t)stream;
    DRR_HIP_CHECK(hipEventRecord(B->ev0, st));
    DRR_HIP_CHECK(hipGraphLaunch(B->exec_march, st));
    DRR_HIP_CHECK(hipEventRecord(B->ev1, st));
    if (B->exec_epi) DRR_HIP_CHECK(hipGraphLaunch(B->exec_epi, st));
    DRR_HIP_CHECK(hipEventRecord(B->ev2, st));
    B->launched = true;
    if (prev >= 0 && prev != B->ctx->device) (void)hipSetDevice(prev);
    return DRR_OK;
}

extern "C" int drr_batch_last_march_ms(drr_batch *B, float *ms) {
    if (!B || !ms) return drr_set_error(DRR_ERR_INVALID, "drr_batch_last_march_ms: NULL argument");
    if (!B->launched) return drr_set_error(DRR_ERR_STATE, "drr_batch_last_march_ms: not launched yet");
    DRR_HIP_CHECK(hipEventSynchronize(B->ev2));
    DRR_HIP_CHECK(hipEventElapsedTime(ms, B->ev0, B->ev1));
    return DRR_OK;
}

extern "C" int drr_batch_destroy(drr_batch *B) {
    if (!B) return DRR_OK;
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != B->ctx->device) (void)hipSetDevice(B->ctx->device);
    if (B->launched) (void)hipEventSynchronize(B->ev2);
    const int dev = B->ctx->device;
    batch_free(B);
    if (prev >= 0 && prev != dev) (void)hipSetDevice(prev);
    return DRR_OK;
}
