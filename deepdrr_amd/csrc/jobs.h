// jobs.h -- job tables and the persistent per-XCD tile queue shared by the
// batched forward (batch.hip) and the batched adjoint (adjoint.hip).
//
// A job is one view of one volume; its detector is cut into tiles of 64
// pixels (one wavefront).  (Views over several volumes are jobs of their own
// table, marched by k_batch_multi in batch.hip.)  The tile list is grouped by volume and job and
// split into 8 contiguous segments, one per XCD.  A persistent wave takes
// tiles from its XCD's segment (atomic counter) and steals from the other
// segments once it is empty, so each XCD's L2 serves a coherent range of
// rays and the tail is balanced at tile granularity.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "drr_internal.h"
#include "epilogue.h"

struct JobDev {
    Geo1 g;
    int vol;
    int tiles_x;
    int64_t out_offset;
};

// A job over 2..4 volumes: the literal multi-volume march (priorities, the
// shared label cache) of the whole view, on 8x8-pixel wave tiles.
struct JobMultiDev {
    ViewDev vd;
    int tiles_x;
    int pad;
    int64_t out_offset;
};

struct JobTable {
    int n_jobs = 0, n_vols = 0;
    JobDev *d_jobs = nullptr;   // single-volume jobs (entries of multi-volume jobs unused)
    VolDev *d_vols = nullptr;
    int2 *d_items = nullptr;    // single-volume tiles (job, tile)
    int *d_seg = nullptr;       // 9 segment starts
    int *d_counters = nullptr;  // 8 queue counters
    int64_t n_items = 0, max_npix = 0;
    int grid = 0, block = 0;
    std::vector<ImgSeg> imgs;   // host copy: offset / size of each job's image
    // multi-volume jobs: their own table, tiles grouped by volume count V
    JobMultiDev *d_mjobs = nullptr;
    int2 *d_mitems = nullptr;   // (index into d_mjobs, tile)
    int n_mjobs = 0;
    int64_t m_first[DRR_MAX_VOLUMES + 1] = {}, m_count[DRR_MAX_VOLUMES + 1] = {};
};

// Build (and upload) the job table; host-side validation included.
// drr_jobs_build accepts the wave tiles of the batched march (8x8, 16x4) and
// jobs over several volumes; drr_jobs_build_tiles any tile (the adjoint's
// 16x16 workgroup tiles) and single-volume jobs only.
int drr_jobs_build(drr_ctx *ctx, const drr_volume *const *vols, int n_vols, const drr_job_t *jobs,
                   int n_jobs, const drr_batch_tuning_t &tune, JobTable &t);
int drr_jobs_build_tiles(drr_ctx *ctx, const drr_volume *const *vols, int n_vols, const drr_job_t *jobs,
                         int n_jobs, const drr_batch_tuning_t &tune, JobTable &t);
void drr_jobs_free(JobTable &t);

__global__ void k_jobs_reset(int *counters);

// Persistent tile loop: calls body(job_index, const JobDev&, udx, vdx, valid)
// for every pixel of every tile; `valid` is false for lanes past the image edge
// (all lanes of the wave call body together so it may use wave reductions).
template <int TW, typename Body>
__device__ __forceinline__ void for_each_tile(const JobDev *__restrict__ jobs, const int2 *__restrict__ items,
                                              const int *__restrict__ seg, int *counters, Body body) {
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
        body(j, J, udx, vdx, udx < J.g.width && vdx < J.g.height);
    }
}

// Persistent loop over workgroup-sized items: thread 0 takes the next item
// (same per-XCD queues and stealing as for_each_tile) and the whole block
// receives it through LDS, so every thread sees the same items and the same
// number of iterations (the body may use __syncthreads).  body(j, J, tile).
template <typename Body>
__device__ __forceinline__ void for_each_block_item(const JobDev *__restrict__ jobs, const int2 *__restrict__ items,
                                                    const int *__restrict__ seg, int *counters, Body body) {
    __shared__ int2 s_item;
    __shared__ int s_q, s_exhausted;
    if (threadIdx.x == 0) {
        s_q = blockIdx.x & 7;
        s_exhausted = 0;
    }
    __syncthreads();
    for (;;) {
        if (threadIdx.x == 0) {
            int2 it = make_int2(-1, -1);
            while (s_exhausted < 8) {
                const int q = s_q;
                const int idx = seg[q] + atomicAdd(&counters[q], 1);
                if (idx < seg[q + 1]) {
                    it = items[idx];
                    break;
                }
                s_q = (q + 1) & 7;
                ++s_exhausted;
            }
            s_item = it;
        }
        __syncthreads();
        // the same value in every lane: make it scalar so the job's uniforms
        // are loaded once per wave into scalar registers
        const int j = __builtin_amdgcn_readfirstlane(s_item.x);
        const int tile = __builtin_amdgcn_readfirstlane(s_item.y);
        __syncthreads();  // s_item is rewritten only after every thread has read it
        if (j < 0) break;
        body(j, jobs[j], tile);
    }
}
