// batch.hip -- batched forward projection: hundreds of (volume, pose,
// detector) jobs in one persistent launch, replayed from a hipGraph
// (drr_batch_*) or launched once (drr_project_jobs).
//
// This is the batched form of deepdrr's per-view loop
// (deepdrr/projector/projector.py:679-685, one projectKernel launch per view).
// Scheduling is described in jobs.h.  The per-ray march is the same device
// code as drr_project, so every image is bit-identical to the one-view path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "drr_internal.h"
#include "epilogue.h"
#include "jobs.h"
#include "march.h"

using namespace drr;

__global__ void k_jobs_reset(int *counters) {
    if (threadIdx.x < 8) counters[threadIdx.x] = 0;
}

void drr_jobs_free(JobTable &t) {
    (void)hipFree(t.d_jobs);
    (void)hipFree(t.d_vols);
    (void)hipFree(t.d_items);
    (void)hipFree(t.d_seg);
    (void)hipFree(t.d_counters);
    (void)hipFree(t.d_mjobs);
    (void)hipFree(t.d_mitems);
    t.d_jobs = nullptr;
    t.d_vols = nullptr;
    t.d_items = nullptr;
    t.d_seg = nullptr;
    t.d_counters = nullptr;
    t.d_mjobs = nullptr;
    t.d_mitems = nullptr;
}

static int jobs_build_impl(drr_ctx *ctx, const drr_volume *const *vols, int n_vols, const drr_job_t *jobs,
                           int n_jobs, const drr_batch_tuning_t &tune, bool allow_multi, JobTable &t);

int drr_jobs_build(drr_ctx *ctx, const drr_volume *const *vols, int n_vols, const drr_job_t *jobs,
                   int n_jobs, const drr_batch_tuning_t &tune, JobTable &t) {
    if (!((tune.tile_w == 8 && tune.tile_h == 8) || (tune.tile_w == 16 && tune.tile_h == 4)))
        return drr_set_error(DRR_ERR_INVALID, "jobs: tile must be 8x8 or 16x4");
    if (tune.waves_per_block < 1 || tune.waves_per_block > 4 || tune.blocks_per_cu < 1 || tune.blocks_per_cu > 32)
        return drr_set_error(DRR_ERR_INVALID, "jobs: waves_per_block in [1,4], blocks_per_cu in [1,32]");
    return jobs_build_impl(ctx, vols, n_vols, jobs, n_jobs, tune, true, t);
}

namespace {
// Host half of a job table: validated jobs, volume descriptors, the tile list
// grouped by volume and job, and the 8 queue segments.
struct JobHost {
    std::vector<JobDev> jobs;
    std::vector<VolDev> vols;
    std::vector<int2> items;
    int seg[9];
    std::vector<JobMultiDev> mjobs;
    std::vector<int2> mitems;
};

// Validate the jobs and fill `h` and the host fields of `t` (n_jobs, n_vols,
// imgs, max_npix, n_items, grid, block, the multi-volume groups).  A plan
// with more than `max_items` ray tiles is refused before its tile list is
// built.  A job over V > 1 volumes (allow_multi) reads vols[volume_index ..
// volume_index + V - 1] as its view.vol[0 .. V-1].
int jobs_prepare(drr_ctx *ctx, const drr_volume *const *vols, int n_vols, const drr_job_t *jobs, int n_jobs,
                 const drr_batch_tuning_t &tune, int64_t max_items, const char *who, bool allow_multi, JobTable &t,
                 JobHost &h) {
    if (n_vols < 1 || n_jobs < 1 || n_jobs > 65535)
        return drr_set_error(DRR_ERR_INVALID, std::string(who) + ": need >= 1 volume and 1..65535 jobs");
    if (tune.tile_w < 1 || tune.tile_h < 1 || tune.waves_per_block < 1 || tune.blocks_per_cu < 1)
        return drr_set_error(DRR_ERR_INVALID, std::string(who) + ": bad tiling");
    for (int v = 0; v < n_vols; v++)
        if (!vols[v] || vols[v]->ctx != ctx)
            return drr_set_error(DRR_ERR_INVALID, std::string(who) + ": NULL or foreign volume");
    t.n_jobs = n_jobs;
    t.n_vols = n_vols;
    h.jobs.assign(n_jobs, JobDev{});
    t.imgs.assign(n_jobs, ImgSeg{0, 0});
    std::vector<int> order(n_jobs);
    t.max_npix = 0;
    int64_t n_items = 0, n_mitems = 0;
    std::vector<int> mjob_of(n_jobs, -1);
    h.mjobs.clear();
    int n_single = 0;
    for (int j = 0; j < n_jobs; j++) {
        const drr_job_t &J = jobs[j];
        const int V = J.view.n_volumes;
        if (V < 1 || V > DRR_MAX_VOLUMES)
            return drr_set_error(DRR_ERR_INVALID, std::string(who) + ": view.n_volumes must be in [1, 4]");
        if (V != 1 && !allow_multi)
            return drr_set_error(DRR_ERR_UNSUPPORTED, std::string(who) + ": each job must view exactly one volume");
        if (J.volume_index < 0 || J.volume_index + V > n_vols)
            return drr_set_error(DRR_ERR_INVALID, std::string(who) + ": volume_index out of range");
        if (J.view.width < 1 || J.view.height < 1 || J.view.width > 65536 || J.view.height > 65536 || J.out_offset < 0)
            return drr_set_error(DRR_ERR_INVALID, std::string(who) + ": bad image size or offset");
        if (drr_check_view_finite(&J.view) != DRR_OK) return DRR_ERR_INVALID;
        t.imgs[j] = ImgSeg{J.out_offset, (int64_t)J.view.width * J.view.height};
        t.max_npix = std::max<int64_t>(t.max_npix, t.imgs[j].npix);
        if (V > 1) {
            JobMultiDev md{};
            view_to_dev(J.view, vols + J.volume_index, md.vd);
            md.tiles_x = (J.view.width + 7) / 8;
            md.out_offset = J.out_offset;
            mjob_of[j] = (int)h.mjobs.size();
            h.mjobs.push_back(md);
            n_mitems += (int64_t)md.tiles_x * ((J.view.height + 7) / 8);
            continue;
        }
        view_to_geo1(J.view, h.jobs[j].g);
        h.jobs[j].vol = J.volume_index;
        h.jobs[j].tiles_x = (J.view.width + tune.tile_w - 1) / tune.tile_w;
        h.jobs[j].out_offset = J.out_offset;
        n_items += (int64_t)h.jobs[j].tiles_x * ((J.view.height + tune.tile_h - 1) / tune.tile_h);
        order[n_single++] = j;
    }
    order.resize(n_single);
    if (n_items + n_mitems >= ((int64_t)1 << 31))
        return drr_set_error(DRR_ERR_INVALID, std::string(who) + ": too many ray tiles");
    if (max_items >= 0 && n_items + n_mitems > max_items)
        return drr_set_error(DRR_ERR_INVALID, std::string(who) + ": the batch has " + std::to_string(n_items + n_mitems) +
                                                  " ray tiles, more than the " + std::to_string(max_items) +
                                                  " its scratch holds");
    // group tiles by volume (L2 / MALL reuse across poses), then by job
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return h.jobs[a].vol < h.jobs[b].vol; });
    h.items.clear();
    h.items.reserve((size_t)n_items);
    for (int j : order) {
        const int ty = (jobs[j].view.height + tune.tile_h - 1) / tune.tile_h;
        const int n = h.jobs[j].tiles_x * ty;
        for (int k = 0; k < n; k++) h.items.push_back(make_int2(j, k));
    }
    t.n_items = n_items;
    // multi-volume tiles, grouped by volume count (one kernel instance each)
    h.mitems.clear();
    h.mitems.reserve((size_t)n_mitems);
    for (int V = 0; V <= DRR_MAX_VOLUMES; V++) {
        t.m_first[V] = (int64_t)h.mitems.size();
        for (int j = 0; j < n_jobs; j++) {
            if (mjob_of[j] < 0 || jobs[j].view.n_volumes != V) continue;
            const JobMultiDev &md = h.mjobs[mjob_of[j]];
            const int n = md.tiles_x * ((jobs[j].view.height + 7) / 8);
            for (int k = 0; k < n; k++) h.mitems.push_back(make_int2(mjob_of[j], k));
        }
        t.m_count[V] = (int64_t)h.mitems.size() - t.m_first[V];
    }
    t.n_mjobs = (int)h.mjobs.size();
    for (int q = 0; q <= 8; q++) h.seg[q] = (int)(t.n_items * q / 8);
    h.vols.resize(n_vols);
    for (int v = 0; v < n_vols; v++) h.vols[v] = vol_to_dev(vols[v]);
    int cus = 0;
    hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
    if (e != hipSuccess) return drr_hip_error(e, "hipDeviceGetAttribute(multiprocessor count)");
    t.block = 64 * tune.waves_per_block;
    t.grid = cus * tune.blocks_per_cu;
    return DRR_OK;
}
}  // namespace

int drr_jobs_build_tiles(drr_ctx *ctx, const drr_volume *const *vols, int n_vols, const drr_job_t *jobs,
                         int n_jobs, const drr_batch_tuning_t &tune, JobTable &t) {
    return jobs_build_impl(ctx, vols, n_vols, jobs, n_jobs, tune, false, t);
}

static int jobs_build_impl(drr_ctx *ctx, const drr_volume *const *vols, int n_vols, const drr_job_t *jobs,
                           int n_jobs, const drr_batch_tuning_t &tune, bool allow_multi, JobTable &t) {
    JobHost h;
    int rc = jobs_prepare(ctx, vols, n_vols, jobs, n_jobs, tune, -1, "jobs", allow_multi, t, h);
    if (rc != DRR_OK) return rc;
    hipError_t e;
#define JCHK(x)                                             \
    do {                                                    \
        e = (x);                                            \
        if (e != hipSuccess) {                              \
            drr_jobs_free(t);                               \
            return drr_hip_error(e, #x);                    \
        }                                                   \
    } while (0)
    JCHK(hipMalloc(&t.d_jobs, sizeof(JobDev) * n_jobs));
    JCHK(hipMalloc(&t.d_vols, sizeof(VolDev) * n_vols));
    JCHK(hipMalloc(&t.d_items, sizeof(int2) * std::max<size_t>(1, h.items.size())));
    JCHK(hipMalloc(&t.d_seg, sizeof(int) * 9));
    JCHK(hipMalloc(&t.d_counters, sizeof(int) * 8));
    JCHK(hipMemcpy(t.d_jobs, h.jobs.data(), sizeof(JobDev) * n_jobs, hipMemcpyHostToDevice));
    JCHK(hipMemcpy(t.d_vols, h.vols.data(), sizeof(VolDev) * n_vols, hipMemcpyHostToDevice));
    if (!h.items.empty())
        JCHK(hipMemcpy(t.d_items, h.items.data(), sizeof(int2) * h.items.size(), hipMemcpyHostToDevice));
    JCHK(hipMemcpy(t.d_seg, h.seg, sizeof(int) * 9, hipMemcpyHostToDevice));
    if (!h.mjobs.empty()) {
        JCHK(hipMalloc(&t.d_mjobs, sizeof(JobMultiDev) * h.mjobs.size()));
        JCHK(hipMalloc(&t.d_mitems, sizeof(int2) * h.mitems.size()));
        JCHK(hipMemcpy(t.d_mjobs, h.mjobs.data(), sizeof(JobMultiDev) * h.mjobs.size(), hipMemcpyHostToDevice));
        JCHK(hipMemcpy(t.d_mitems, h.mitems.data(), sizeof(int2) * h.mitems.size(), hipMemcpyHostToDevice));
    }
#undef JCHK
    return DRR_OK;
}

struct drr_batch_s {
    drr_ctx *ctx = nullptr;
    int M = 0;
    bool atten = false;
    bool fast = false;  // DRR_MARCH_FAST at creation
    drr_batch_tuning_t tune{};
    JobTable t;
    ImgSeg *d_imgs = nullptr;
    unsigned int *d_keys = nullptr;
    hipStream_t cap = nullptr;
    hipGraph_t graph_march = nullptr, graph_epi = nullptr;
    hipGraphExec_t exec_march = nullptr, exec_epi = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;  // before march, after march, after epilogue
    bool launched = false;
};

template <int M, bool ATTEN, int TW, bool FAST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MarchOcc<M + (ATTEN ? 1 : 0)>::waves)))
void k_batch_march(const JobDev *__restrict__ jobs, const VolDev *__restrict__ vols,
                   const int2 *__restrict__ items, const int *__restrict__ seg, int *counters,
                   SpectrumDev sp, float step, int air, float *__restrict__ I, float *__restrict__ P,
                   float *__restrict__ AD) {
    for_each_tile<TW>(jobs, items, seg, counters, [&](int, const JobDev &J, int udx, int vdx, bool valid) {
        if (!valid) return;
        float ad[M];
        march_v1<M, ATTEN, FAST>(J.g, vols[J.vol], udx, vdx, step, air, ad);
        float inten, pp;
        spectral<M>(sp, ad, inten, pp);
        const int64_t pix = (int64_t)vdx * J.g.width + udx;
        const int64_t o = J.out_offset + pix;
        I[o] = inten;
        if (P) P[o] = pp;
        if (AD) {
            // job block of M planes at M * out_offset (jobs' images do not overlap)
            const int64_t npix = (int64_t)J.g.width * J.g.height;
#pragma unroll
            for (int m = 0; m < M; m++) AD[M * J.out_offset + m * npix + pix] = ad[m];
        }
    });
}

template <int M, bool ATTEN, bool FAST>
static void launch_march_tw(int tw, const JobTable &t, hipStream_t st, SpectrumDev sp, float step, int air,
                            float *I, float *P, float *AD) {
    dim3 g(t.grid), b(t.block);
    if (tw == 16)
        k_batch_march<M, ATTEN, 16, FAST><<<g, b, 0, st>>>(t.d_jobs, t.d_vols, t.d_items, t.d_seg, t.d_counters,
                                                            sp, step, air, I, P, AD);
    else
        k_batch_march<M, ATTEN, 8, FAST><<<g, b, 0, st>>>(t.d_jobs, t.d_vols, t.d_items, t.d_seg, t.d_counters,
                                                           sp, step, air, I, P, AD);
}

template <int M>
static void launch_march_m(bool atten, bool fast, int tw, const JobTable &t, hipStream_t st, SpectrumDev sp,
                           float step, int air, float *I, float *P, float *AD) {
    if (fast) {
        if (atten)
            launch_march_tw<M, true, true>(tw, t, st, sp, step, air, I, P, AD);
        else
            launch_march_tw<M, false, true>(tw, t, st, sp, step, air, I, P, AD);
    } else {
        if (atten)
            launch_march_tw<M, true, false>(tw, t, st, sp, step, air, I, P, AD);
        else
            launch_march_tw<M, false, false>(tw, t, st, sp, step, air, I, P, AD);
    }
}

static void launch_march(const JobTable &t, int M, bool atten, bool fast, int tw, hipStream_t st,
                         SpectrumDev sp, float step, int air, float *I, float *P, float *AD) {
    switch (M) {
#define DRR_CASE(MM)                                                             \
    case MM:                                                                     \
        launch_march_m<MM>(atten, fast, tw, t, st, sp, step, air, I, P, AD);     \
        break;
        DRR_CASE(1) DRR_CASE(2) DRR_CASE(3) DRR_CASE(4) DRR_CASE(5) DRR_CASE(6) DRR_CASE(7) DRR_CASE(8) \
        DRR_CASE(9) DRR_CASE(10) DRR_CASE(11) DRR_CASE(12) DRR_CASE(13) DRR_CASE(14) DRR_CASE(15) DRR_CASE(16)
#undef DRR_CASE
    }
}

// Jobs over 2..4 volumes: one 8x8-pixel tile per wave, the literal
// multi-volume march (march_multi, as drr_project runs such a view).
template <int V, int M>
__global__ __launch_bounds__(256) void k_batch_multi(const JobMultiDev *__restrict__ jobs, const int2 *__restrict__ items,
                                                     int64_t n_items, SpectrumDev sp, float step, int atten, int air,
                                                     float *__restrict__ I, float *__restrict__ P,
                                                     float *__restrict__ AD) {
    const int64_t it = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (it >= n_items) return;  // whole waves leave together
    const int lane = threadIdx.x & 63;
    const int2 item = items[it];
    const int j = __builtin_amdgcn_readfirstlane(item.x);
    const int tile = __builtin_amdgcn_readfirstlane(item.y);
    const JobMultiDev &J = jobs[j];
    int udx, vdx;
    lane_pixel(tile % J.tiles_x, tile / J.tiles_x, lane, udx, vdx);
    if (udx >= J.vd.width || vdx >= J.vd.height) return;
    float ad[M];
    march_multi<V, M>(J.vd, udx, vdx, step, atten, air, ad);
    float inten, pp;
    spectral<M>(sp, ad, inten, pp);
    const int64_t npix = (int64_t)J.vd.width * J.vd.height;
    const int64_t pix = (int64_t)vdx * J.vd.width + udx;
    I[J.out_offset + pix] = inten;
    if (P) P[J.out_offset + pix] = pp;
    if (AD) {
#pragma unroll
        for (int m = 0; m < M; m++) AD[M * J.out_offset + m * npix + pix] = ad[m];
    }
}

template <int V>
static void launch_multi_v(const JobTable &t, int M, hipStream_t st, SpectrumDev sp, float step, int atten, int air,
                           float *I, float *P, float *AD) {
    const int64_t n = t.m_count[V];
    if (n <= 0) return;
    const dim3 g((unsigned)((n + 3) / 4)), b(256);
    const int2 *items = t.d_mitems + t.m_first[V];
    switch (M) {
#define DRR_CASE(MM) \
    case MM: k_batch_multi<V, MM><<<g, b, 0, st>>>(t.d_mjobs, items, n, sp, step, atten, air, I, P, AD); break;
        DRR_CASE(1) DRR_CASE(2) DRR_CASE(3) DRR_CASE(4) DRR_CASE(5) DRR_CASE(6) DRR_CASE(7) DRR_CASE(8) \
        DRR_CASE(9) DRR_CASE(10) DRR_CASE(11) DRR_CASE(12) DRR_CASE(13) DRR_CASE(14) DRR_CASE(15) DRR_CASE(16)
#undef DRR_CASE
    }
}

// Every job of a table: the persistent single-volume march, then the
// multi-volume groups.
static void launch_all(const JobTable &t, int M, bool atten, bool fast, int tw, hipStream_t st, SpectrumDev sp,
                       float step, int air, float *I, float *P, float *AD) {
    if (t.n_items > 0) launch_march(t, M, atten, fast, tw, st, sp, step, air, I, P, AD);
    launch_multi_v<2>(t, M, st, sp, step, atten ? 1 : 0, air, I, P, AD);
    launch_multi_v<3>(t, M, st, sp, step, atten ? 1 : 0, air, I, P, AD);
    launch_multi_v<4>(t, M, st, sp, step, atten ? 1 : 0, air, I, P, AD);
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
    drr_jobs_free(B->t);
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
    k_jobs_reset<<<1, 64, 0, B->cap>>>(B->t.d_counters);
    launch_all(B->t, B->M, B->atten, B->fast, B->tune.tile_w, B->cap, sp, step, air, I, P, nullptr);
    e = hipStreamEndCapture(B->cap, &B->graph_march);
    if (e != hipSuccess) return e;
    e = hipGraphInstantiate(&B->exec_march, B->graph_march, nullptr, nullptr, 0);
    if (e != hipSuccess || !neglog) return e;
    e = hipStreamBeginCapture(B->cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    int rc = drr_enqueue_neglog(I, images, B->d_imgs, B->t.n_jobs, B->t.max_npix, B->d_keys, 0.01f, B->cap);
    hipError_t e2 = hipStreamEndCapture(B->cap, &B->graph_epi);
    if (rc != DRR_OK) return hipErrorLaunchFailure;
    if (e2 != hipSuccess) return e2;
    return hipGraphInstantiate(&B->exec_epi, B->graph_epi, nullptr, nullptr, 0);
}

namespace {
struct DevSwitch {
    int prev = -1, dev;
    explicit DevSwitch(int d) : dev(d) {
        (void)hipGetDevice(&prev);
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DevSwitch() {
        if (prev >= 0 && prev != dev) (void)hipSetDevice(prev);
    }
};
}  // namespace

extern "C" int drr_batch_create(drr_ctx *ctx, const drr_volume *const *vols, int n_vols,
                                const drr_job_t *jobs, int n_jobs, const drr_batch_tuning_t *tuning,
                                float *intensity, float *photon_prob, float *images, int neglog,
                                drr_batch **out) {
    if (!ctx || !vols || !jobs || !out || !intensity || !photon_prob)
        return drr_set_error(DRR_ERR_INVALID, "drr_batch_create: NULL argument");
    *out = nullptr;
    if (neglog && !images) return drr_set_error(DRR_ERR_INVALID, "drr_batch_create: neglog needs images");
    if (!ctx->have_scene || ctx->n_bins < 1)
        return drr_set_error(DRR_ERR_STATE, "drr_batch_create: scene/spectrum not set");
    DevSwitch ds(ctx->device);
    drr_batch *B = new drr_batch();
    B->ctx = ctx;
    B->M = ctx->scene.n_materials;
    B->atten = ctx->scene.attenuate_outside != 0;
    B->fast = ctx->scene.march == DRR_MARCH_FAST;
    B->tune = tuning ? *tuning : drr_batch_tuning_t{8, 8, 4, 8};
    int rc = drr_jobs_build(ctx, vols, n_vols, jobs, n_jobs, B->tune, B->t);
    if (rc != DRR_OK) {
        batch_free(B);
        return rc;
    }
    hipError_t e;
#define BCHK(x)                                                        \
    do {                                                               \
        e = (x);                                                       \
        if (e != hipSuccess) {                                         \
            batch_free(B);                                             \
            return drr_hip_error(e, #x);                               \
        }                                                              \
    } while (0)
    BCHK(hipMalloc(&B->d_imgs, sizeof(ImgSeg) * n_jobs));
    BCHK(hipMalloc(&B->d_keys, sizeof(unsigned int) * 4 * n_jobs));
    BCHK(hipMemcpy(B->d_imgs, B->t.imgs.data(), sizeof(ImgSeg) * n_jobs, hipMemcpyHostToDevice));
    BCHK(hipStreamCreateWithFlags(&B->cap, hipStreamNonBlocking));
    BCHK(hipEventCreate(&B->ev0));
    BCHK(hipEventCreate(&B->ev1));
    BCHK(hipEventCreate(&B->ev2));
    SpectrumDev sp{ctx->d_energies, ctx->d_pdf, ctx->d_mu, ctx->n_bins};
    BCHK(capture(B, sp, ctx->scene.step, ctx->scene.air_index, intensity, photon_prob, images, neglog));
#undef BCHK
    *out = B;
    return DRR_OK;
}

extern "C" int drr_batch_launch(drr_batch *B, void *stream) {
    if (!B) return drr_set_error(DRR_ERR_INVALID, "drr_batch_launch: NULL batch");
    DevSwitch ds(B->ctx->device);
    hipStream_t st = (hipStream_t)stream;
    DRR_HIP_CHECK(hipEventRecord(B->ev0, st));
    DRR_HIP_CHECK(hipGraphLaunch(B->exec_march, st));
    DRR_HIP_CHECK(hipEventRecord(B->ev1, st));
    if (B->exec_epi) DRR_HIP_CHECK(hipGraphLaunch(B->exec_epi, st));
    DRR_HIP_CHECK(hipEventRecord(B->ev2, st));
    B->launched = true;
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
    DevSwitch ds(B->ctx->device);
    if (B->launched) (void)hipEventSynchronize(B->ev2);
    batch_free(B);
    return DRR_OK;
}

extern "C" int drr_project_jobs(drr_ctx *ctx, const drr_volume *const *vols, int n_vols, const drr_job_t *jobs,
                                int n_jobs, float *intensity, float *photon_prob, float *area_density,
                                void *stream) {
    if (!ctx || !vols || !jobs || !intensity)
        return drr_set_error(DRR_ERR_INVALID, "drr_project_jobs: NULL argument");
    if (!ctx->have_scene || ctx->n_bins < 1)
        return drr_set_error(DRR_ERR_STATE, "drr_project_jobs: scene/spectrum not set");
    DevSwitch ds(ctx->device);
    JobTable t;
    const drr_batch_tuning_t tune{8, 8, 4, 8};
    int rc = drr_jobs_build(ctx, vols, n_vols, jobs, n_jobs, tune, t);
    if (rc != DRR_OK) return rc;
    hipStream_t st = (hipStream_t)stream;
    SpectrumDev sp{ctx->d_energies, ctx->d_pdf, ctx->d_mu, ctx->n_bins};
    (void)hipEventRecord(ctx->ev0, st);
    k_jobs_reset<<<1, 64, 0, st>>>(t.d_counters);
    launch_all(t, ctx->scene.n_materials, ctx->scene.attenuate_outside != 0, ctx->scene.march == DRR_MARCH_FAST,
               tune.tile_w, st, sp, ctx->scene.step, ctx->scene.air_index, intensity, photon_prob, area_density);
    hipError_t e = hipGetLastError();
    (void)hipEventRecord(ctx->ev1, st);
    ctx->timed = true;
    hipError_t e2 = hipStreamSynchronize(st);
    drr_jobs_free(t);
    if (e != hipSuccess) return drr_hip_error(e, "drr_project_jobs launch");
    if (e2 != hipSuccess) return drr_hip_error(e2, "drr_project_jobs");
    return DRR_OK;
}

// ---- drr_project_jobs without device allocation ------------------------------
//
// The job table, volume descriptors, queue segments and counters and the tile
// list of one call live in `scratch`, a file-scope device array of
// DRR_NOALLOC_SCRATCH_BYTES that every device holds until the process ends.
// In host code `scratch` names a host-side stand-in, so its device address is
// looked up once per device (hipGetSymbolAddress, with that device current)
// and only that address is handed to copies and kernels.  All calls on one
// device share the array: calls that overlap in time on one device race.

static __device__ __align__(256) unsigned char scratch[DRR_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (the current device).
int scratch_address(int device, unsigned char **out) {
    if (device < 0 || device >= kMaxDevices)
        return drr_set_error(DRR_ERR_UNSUPPORTED, "drr_project_jobs_noalloc: device id out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return drr_hip_error(e, "hipGetSymbolAddress(scratch)");
        if (!p) return drr_set_error(DRR_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[device] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[device];
    return DRR_OK;
}

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
}  // namespace

extern "C" int drr_project_jobs_noalloc(drr_ctx *ctx, const drr_volume *const *vols, int n_vols,
                                        const drr_job_t *jobs, int n_jobs, float *intensity,
                                        float *photon_prob, float *area_density, void *stream) {
    if (!ctx || !vols || !jobs || !intensity)
        return drr_set_error(DRR_ERR_INVALID, "drr_project_jobs_noalloc: NULL argument");
    if (!ctx->have_scene || ctx->n_bins < 1)
        return drr_set_error(DRR_ERR_STATE, "drr_project_jobs_noalloc: scene/spectrum not set");
    if (n_vols < 1 || n_jobs < 1 || n_jobs > 65535)
        return drr_set_error(DRR_ERR_INVALID, "drr_project_jobs_noalloc: need >= 1 volume and 1..65535 jobs");
    DevSwitch ds(ctx->device);
    unsigned char *base = nullptr;
    int rc = scratch_address(ctx->device, &base);
    if (rc != DRR_OK) return rc;
    // layout: counters | segments | volumes | jobs | multi-volume jobs | tiles
    // (single-volume, then multi-volume), each 256-B aligned; a job over
    // several volumes takes room in both job tables
    int n_multi = 0;
    for (int j = 0; j < n_jobs; j++) n_multi += jobs[j].view.n_volumes > 1 ? 1 : 0;
    const size_t off_seg = 256, off_vols = 512;
    const size_t off_jobs = align256(off_vols + sizeof(VolDev) * (size_t)n_vols);
    const size_t off_mjobs = align256(off_jobs + sizeof(JobDev) * (size_t)n_jobs);
    const size_t off_items = align256(off_mjobs + sizeof(JobMultiDev) * (size_t)n_multi);
    if (off_items > DRR_NOALLOC_SCRATCH_BYTES)
        return drr_set_error(DRR_ERR_INVALID, "drr_project_jobs_noalloc: the job table exceeds the " +
                                                  std::to_string(DRR_NOALLOC_SCRATCH_BYTES) + "-byte scratch");
    const int64_t max_items = (int64_t)((DRR_NOALLOC_SCRATCH_BYTES - off_items) / sizeof(int2));
    JobTable t;
    JobHost h;
    const drr_batch_tuning_t tune{8, 8, 4, 8};
    rc = jobs_prepare(ctx, vols, n_vols, jobs, n_jobs, tune, max_items, "drr_project_jobs_noalloc", true, t, h);
    if (rc != DRR_OK) return rc;
    t.d_counters = (int *)base;
    t.d_seg = (int *)(base + off_seg);
    t.d_vols = (VolDev *)(base + off_vols);
    t.d_jobs = (JobDev *)(base + off_jobs);
    t.d_mjobs = (JobMultiDev *)(base + off_mjobs);
    t.d_items = (int2 *)(base + off_items);
    t.d_mitems = t.d_items + h.items.size();
    hipStream_t st = (hipStream_t)stream;
    // The host tables must outlive the copies that read them: every return
    // after the first copy is queued synchronizes the stream first.
    hipError_t e;
#define NCHK(x)                                      \
    do {                                             \
        e = (x);                                     \
        if (e != hipSuccess) {                       \
            (void)hipStreamSynchronize(st);          \
            return drr_hip_error(e, #x);             \
        }                                            \
    } while (0)
    NCHK(hipMemcpyAsync(t.d_seg, h.seg, sizeof(int) * 9, hipMemcpyHostToDevice, st));
    NCHK(hipMemcpyAsync(t.d_vols, h.vols.data(), sizeof(VolDev) * n_vols, hipMemcpyHostToDevice, st));
    NCHK(hipMemcpyAsync(t.d_jobs, h.jobs.data(), sizeof(JobDev) * n_jobs, hipMemcpyHostToDevice, st));
    if (!h.items.empty())
        NCHK(hipMemcpyAsync(t.d_items, h.items.data(), sizeof(int2) * h.items.size(), hipMemcpyHostToDevice, st));
    if (!h.mjobs.empty()) {
        NCHK(hipMemcpyAsync(t.d_mjobs, h.mjobs.data(), sizeof(JobMultiDev) * h.mjobs.size(), hipMemcpyHostToDevice, st));
        NCHK(hipMemcpyAsync(t.d_mitems, h.mitems.data(), sizeof(int2) * h.mitems.size(), hipMemcpyHostToDevice, st));
    }
#undef NCHK
    SpectrumDev sp{ctx->d_energies, ctx->d_pdf, ctx->d_mu, ctx->n_bins};
    (void)hipEventRecord(ctx->ev0, st);
    k_jobs_reset<<<1, 64, 0, st>>>(t.d_counters);  // the previous call's counters are left in scratch
    launch_all(t, ctx->scene.n_materials, ctx->scene.attenuate_outside != 0, ctx->scene.march == DRR_MARCH_FAST,
               tune.tile_w, st, sp, ctx->scene.step, ctx->scene.air_index, intensity, photon_prob, area_density);
    e = hipGetLastError();
    (void)hipEventRecord(ctx->ev1, st);
    ctx->timed = true;
    hipError_t e2 = hipStreamSynchronize(st);
    if (e != hipSuccess) return drr_hip_error(e, "drr_project_jobs_noalloc launch");
    if (e2 != hipSuccess) return drr_hip_error(e2, "drr_project_jobs_noalloc");
    return DRR_OK;
}
