// drr_internal.h -- shared definitions of the HIP library (not installed).
//
// Volume layout in HBM ("cell-expanded"): for a volume of nx*ny*nz voxels the
// library stores one 40-byte record per trilinear cell (i, j, k), i in
// [-1, nx-1] (same for j, k), i.e. (nx+1)(ny+1)(nz+1) cells, i fastest:
//   words 0-7   : 8 float32 corner densities, corner d = dx + 2 dy + 4 dz
//                 holding D[clamp(i+dx)][clamp(j+dy)][clamp(k+dz)]   (32 B)
//   word 8      : the 8 corner labels as 4-bit fields, corner d at bits
//                 4d..4d+3 (labels are < DRR_MAX_MATERIALS = 16)
//   word 9      : their common label if all 8 agree, else ~0          (8 B)
// A ray sample with base = floor(p) reads cell clamp(base, -1, n-1): three
// loads from one address (16 + 16 + 8 bytes) replace the 16 scattered texel
// fetches of the reference's texture path (project_kernel.cu:421-425, 542).
// Clamping the cell index reproduces clamp-to-edge addressing exactly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/deepdrr_amd.h"

struct drr_volume_s {
    drr_ctx *ctx = nullptr;
    int device = 0;  // HIP device of the allocation (outlives the context)
    int nx = 0, ny = 0, nz = 0;
    int64_t n_cells = 0;
    uint32_t *cells = nullptr;  // DRR_CELL_WORDS 32-bit words per cell
    // voxel-order records [nz][ny][nx] of (density bits, label), made on first
    // use by the Monte-Carlo path: a nearest-voxel lookup is one 8-byte load
    // (one memory sector, where separate density and label arrays take two)
    uint2 *lin_vox = nullptr;
    float max_density = 0.0f;
};

struct drr_ctx_s {
    int device = 0;
    drr_scene_t scene{0.1f, 0, 0, 0, DRR_MARCH_EXACT};
    bool have_scene = false;
    int n_bins = 0;
    float *d_energies = nullptr;  // [B]
    float *d_pdf = nullptr;       // [B]
    float *d_mu = nullptr;        // [B * M]
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
};

// ---- error plumbing (defined in runtime.hip) ------------------------------
int drr_set_error(int code, const std::string &msg);
int drr_hip_error(hipError_t e, const char *what);

#define DRR_HIP_CHECK(expr)                                   \
    do {                                                      \
        hipError_t _e = (expr);                               \
        if (_e != hipSuccess) return drr_hip_error(_e, #expr); \
    } while (0)

// ---- kernel parameter blocks ------------------------------------------------

#define DRR_CELL_WORDS 10  // words of one cell record: 8 densities, 2 label words

// One volume as the march sees it.
struct VolDev {
    const uint32_t *cells;
    int nx, ny, nz;
    // Cell index in FP32 arithmetic for clamped cell coordinates (cx, cy, cz)
    // (march.h, Cell::load): with e1 = cz * a1 + (cy * b1 + c1) and
    // e2 = cy * b2 + (cx + c2) (parameters from vol_index_params),
    //   (uint32)e1 * s1 + bits(e2) = index + 0x4B000000,
    // every intermediate an exact integer below 2^24.
    float a1, b1, c1, b2, c2;
    uint32_t s1;
    float cmax[3];  // nx - 1, ny - 1, nz - 1: the largest cell coordinate
    uint64_t cells_b;  // address of the record of index -0x4B000000
};

// Parameters of the FP32 cell index.  Form A ((nx+1)(ny+1) <= 2^23):
// e1 = cz + 1, e2 = (cy+1)(nx+1) + (cx+1) + 2^23, s1 = (nx+1)(ny+1).  Form B
// ((nz+1)(ny+1) <= 2^23): e1 = (cz+1)(ny+1) + (cy+1), e2 = cx + 1 + 2^23,
// s1 = nx+1.  Dimensions <= 16384 and fewer than 2^31 cells leave no volume
// outside both forms ((nx+1)(ny+1) > 2^23 forces nz+1 < 256).  bits(e2) of
// an integer e2 in [2^23, 2^24) is 0x4B000000 + (e2 - 2^23); the bias is
// taken off the base addresses instead of each index.
inline VolDev vol_index_params(const uint32_t *cells, int nx, int ny, int nz) {
    VolDev d{};
    d.cells = cells;
    d.nx = nx;
    d.ny = ny;
    d.nz = nz;
    d.cmax[0] = (float)(nx - 1);
    d.cmax[1] = (float)(ny - 1);
    d.cmax[2] = (float)(nz - 1);
    const double nx1 = nx + 1.0, ny1 = ny + 1.0, nz1 = nz + 1.0;
    const double two23 = 8388608.0;
    if (nx1 * ny1 <= two23) {
        d.a1 = 1.0f;
        d.b1 = 0.0f;
        d.c1 = 1.0f;
        d.b2 = (float)nx1;
        d.c2 = (float)(nx1 + 1.0 + two23);
        d.s1 = (uint32_t)(nx1 * ny1);
    } else {  // then nz + 1 < 256, so (nz+1)(ny+1) < 2^23
        d.a1 = (float)ny1;
        d.b1 = 1.0f;
        d.c1 = (float)(ny1 + 1.0);
        d.b2 = 0.0f;
        d.c2 = (float)(1.0 + two23);
        d.s1 = (uint32_t)nx1;
        (void)nz1;
    }
    d.cells_b = (uint64_t)(uintptr_t)cells - (uint64_t)0x4B000000u * (4u * DRR_CELL_WORDS);
    return d;
}

// Everything the per-ray march needs for one view of up to V volumes.
struct ViewDev {
    float W[9];
    float max_ray_length;
    int width, height;
    float A[DRR_MAX_VOLUMES][12];
    float s[DRR_MAX_VOLUMES][3];
    float lo[DRR_MAX_VOLUMES][3];
    float hi[DRR_MAX_VOLUMES][3];
    int prio[DRR_MAX_VOLUMES];
    int en[DRR_MAX_VOLUMES];
    VolDev vol[DRR_MAX_VOLUMES];
};

// One view of a single volume: the uniforms of the single-volume march.
struct Geo1 {
    float W[9];
    float mrl;
    int width, height;
    float A[12];
    float s[3];
    float lo[3];
    float hi[3];
    int en;
};

inline void view_to_geo1(const drr_view_t &v, Geo1 &g) {
    for (int i = 0; i < 9; i++) g.W[i] = v.world_from_index[i];
    g.mrl = v.max_ray_length;
    g.width = v.width;
    g.height = v.height;
    for (int i = 0; i < 12; i++) g.A[i] = v.vol[0].ijk_from_world[i];
    for (int i = 0; i < 3; i++) {
        g.s[i] = v.vol[0].source_ijk[i];
        g.lo[i] = v.vol[0].bounds_lo[i];
        g.hi[i] = v.vol[0].bounds_hi[i];
    }
    g.en = v.vol[0].enabled;
}

struct SpectrumDev {
    const float *energies;
    const float *pdf;
    const float *mu;
    int n_bins;
};

// Index + 0x4B000000 of the cell at clamped coordinates (cx, cy, cz), each an
// integer in [-1, n-1] held in FP32 (the march's Cell::load; host code runs it
// too, to check it against integer arithmetic).
__host__ __device__ __forceinline__ uint32_t cell_index_biased(const VolDev &v, float cx, float cy, float cz) {
    const float e1 = __builtin_fmaf(cz, v.a1, __builtin_fmaf(cy, v.b1, v.c1));
    const float e2 = __builtin_fmaf(cy, v.b2, cx + v.c2);
    uint32_t bits;
    __builtin_memcpy(&bits, &e2, 4);
    return (uint32_t)e1 * v.s1 + bits;
}

inline VolDev vol_to_dev(const drr_volume *v) {
    return vol_index_params(v->cells, v->nx, v->ny, v->nz);
}

inline void view_to_dev(const drr_view_t &v, const drr_volume *const *vols, ViewDev &d) {
    for (int i = 0; i < 9; i++) d.W[i] = v.world_from_index[i];
    d.max_ray_length = v.max_ray_length;
    d.width = v.width;
    d.height = v.height;
    for (int k = 0; k < DRR_MAX_VOLUMES; k++) {
        const drr_view_volume_t &vv = v.vol[k];
        for (int i = 0; i < 12; i++) d.A[k][i] = vv.ijk_from_world[i];
        for (int i = 0; i < 3; i++) {
            d.s[k][i] = vv.source_ijk[i];
            d.lo[k][i] = vv.bounds_lo[i];
            d.hi[k][i] = vv.bounds_hi[i];
        }
        d.prio[k] = vv.priority;
        d.en[k] = vv.enabled;
        if (k < v.n_volumes && vols && vols[k]) {
            d.vol[k] = vol_to_dev(vols[k]);
        } else {
            d.vol[k] = vol_index_params(nullptr, 1, 1, 1);
        }
    }
}

// Validation shared by the entry points.
int drr_check_view(const drr_ctx *ctx, const drr_volume *const *vols, const drr_view_t *view);
// Every float of the view's geometry is finite (else DRR_ERR_INVALID).
int drr_check_view_finite(const drr_view_t *view);

// Make vol->lin_vox and vol->max_density if absent (runtime.hip).
int drr_volume_ensure_linear(drr_volume *vol);

// Launchers implemented in forward.hip (used by batch/adjoint code).
int drr_launch_expand(const float *density, const uint8_t *labels, int nx, int ny, int nz,
                      uint32_t *cells, hipStream_t st);
