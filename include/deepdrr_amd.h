/*
 * deepdrr_amd.h -- C ABI of the MI355X DRR engine (libdeepdrr_amd.so).
 *
 * Plain pointers and sizes only; every call returns an int status
 * (DRR_OK = 0, negative on error) and never throws.  The message of the
 * last failure on the calling thread is returned by drr_last_error().
 *
 * Replaces the reference's CUDA module interface:
 *   - projectKernel's 40 launch arguments + 7 -D macros
 *       (deepdrr/projector/project_kernel.cu:135-181,
 *        deepdrr/projector/projector.py:323-392, 718-774)
 *   - texture objects for density (linear, clamp) and labels (point, clamp)
 *       (deepdrr/projector/projector.py:116-257, 1464-1539)
 *   - the per-view uniform uploads (deepdrr/projector/projector.py:802-831)
 *
 * Device buffers passed in (intensity, photon_prob, gradients, images) are
 * HIP device pointers; `stream` is a hipStream_t (NULL = default stream).
 */
#ifndef DEEPDRR_AMD_H
#define DEEPDRR_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DRR_OK 0
#define DRR_ERR_INVALID (-1)
#define DRR_ERR_HIP (-2)
#define DRR_ERR_ALLOC (-3)
#define DRR_ERR_UNSUPPORTED (-4)
#define DRR_ERR_STATE (-5)

#define DRR_MAX_VOLUMES 4
#define DRR_MAX_MATERIALS 16
#define DRR_MAX_BINS 256

typedef struct drr_ctx_s drr_ctx;
typedef struct drr_volume_s drr_volume;
typedef struct drr_batch_s drr_batch;
typedef struct drr_mesh_s drr_mesh;

/* Per-volume part of one view (projector.py:802-831, 1626-1637). */
typedef struct {
    float ijk_from_world[12]; /* 3x4 row-major, stride 12 (project_kernel.cu:273-279) */
    float source_ijk[3];      /* source point in this volume's IJK frame */
    float bounds_lo[3];       /* (-0.5, -0.5, -0.5) */
    float bounds_hi[3];       /* (nx - 0.5, ny - 0.5, nz - 0.5) */
    int32_t priority;         /* lower value wins (project_kernel.cu:488-495) */
    int32_t enabled;          /* 0 = skipped (project_kernel.cu:266-270) */
} drr_view_volume_t;

/* One view: camera rays and every volume's pose in it. */
typedef struct {
    float world_from_index[9]; /* rows 0-2 of the 4x3 world_from_index (projector.py:803) */
    float max_ray_length;      /* <= 0 means unbounded (project_kernel.cu:285) */
    int32_t width;             /* detector pixels along u */
    int32_t height;            /* detector pixels along v */
    int32_t n_volumes;
    drr_view_volume_t vol[DRR_MAX_VOLUMES];
} drr_view_t;

/* How a single-volume view is marched (drr_scene_t.march).  EXACT: the
 * reference's per-sample trilinear filter with its FP32 operation order
 * (area densities bit-identical to the oracle's transcription).  FAST: the same
 * samples, each term still added into its material's area density one sample
 * at a time, with each sample's density formed from a per-cell cubic along the
 * ray (a cell with two labels: nested lerps at the sample's position; with
 * more: the exact filter), so values are within rounding of EXACT and images
 * within 1e-5 relative.  Views over several volumes or with meshes are marched
 * exactly in both modes.  No counterpart in the reference. */
#define DRR_MARCH_EXACT 0
#define DRR_MARCH_FAST 1

/* Scene-wide settings (the reference's compile-time macros and Projector kwargs). */
typedef struct {
    float step;                /* mm along the unit world ray (projector.py:509) */
    int32_t n_materials;       /* NUM_MATERIALS */
    int32_t attenuate_outside; /* ATTENUATE_OUTSIDE_VOLUME */
    int32_t air_index;         /* AIR_INDEX */
    int32_t march;             /* DRR_MARCH_EXACT (0) or DRR_MARCH_FAST */
} drr_scene_t;

/* ---- errors / context ---------------------------------------------------- */

/* Message of the last failed call on this thread ("" if none). */
const char *drr_last_error(void);

/* Number of HIP devices visible (0 without a GPU). */
int drr_device_count(int *count);

/* Create a context bound to HIP device `device_id` (Projector.initialize,
 * projector.py:1395-1421). */
int drr_ctx_create(int device_id, drr_ctx **out);
int drr_ctx_destroy(drr_ctx *ctx);

/* Scene settings: step, material count, air handling (projector.py:365-386). */
int drr_ctx_set_scene(drr_ctx *ctx, const drr_scene_t *scene);

/* Spectrum and mu/rho table, host pointers: energies_keV[B], pdf[B],
 * mu_over_rho[B * n_materials] indexed [bin * M + m] (projector.py:1660-1689). */
int drr_ctx_set_spectrum(drr_ctx *ctx, int n_bins, const float *energies_keV,
                         const float *pdf, const float *mu_over_rho);

/* ---- volumes --------------------------------------------------------------- */

/* Upload one volume.  density: float32 [nz][ny][nx] (i fastest), labels:
 * uint8 [nz][ny][nx] holding global material indices < n_materials.
 * `on_device` = 1 if both pointers are HIP device pointers.  The volume is
 * stored cell-expanded in HBM: for every cell (i, j, k) in [-1, n-1]^3 one
 * 40-byte record of the 8 clamped corner densities (32 B), the 8 corner
 * labels as 4-bit fields (so at most 16 materials) and their common label
 * (8 B).
 * Replaces create_cuda_texture (projector.py:116-257, 1464-1539). */
int drr_volume_create(drr_ctx *ctx, const float *density, const uint8_t *labels,
                      int nx, int ny, int nz, int on_device, drr_volume **out);

/* Upload an int16 HU volume and derive density + labels on the GPU:
 * rho = max(min(0.001029 HU + 1.030, 0.0005886 HU + 1.03), 0) evaluated in
 * FP64 and rounded to float32 (volume.py:338-351); label = label_of_class[c]
 * for the first class c with HU <= upper_hu[c] (load_dicom.py:132-143).
 * hu: int16 [nz][ny][nx], host or device per `on_device`. */
int drr_volume_create_from_hu(drr_ctx *ctx, const int16_t *hu, int nx, int ny, int nz,
                              int on_device, int n_classes, const float *upper_hu,
                              const uint8_t *label_of_class, drr_volume **out);

int drr_volume_destroy(drr_volume *vol);

/* Copy the volume's float32 density and uint8 labels back ([nz][ny][nx], host). */
int drr_volume_download(drr_volume *vol, float *density, uint8_t *labels);

/* ---- forward projection ------------------------------------------------ */

/* One view, one launch: the reference's projectKernel (project_kernel.cu:135-650).
 * vols holds view->n_volumes handles (read up to that count once n_volumes is
 * checked to be in [1, 4]); vols[v] pairs with view->vol[v].  Outputs are device
 * float32 [height][width]
 * (row-major; the reference's column-major buffer after its host swapaxes,
 * projector.py:786-792).  area_density (optional, may be NULL) receives the
 * per-material area densities in g/cm^2 as [n_materials][height][width].
 * solid_angle (optional) receives calculate_solid_angle (project_kernel.cu:14-133). */
int drr_project(drr_ctx *ctx, const drr_volume *const *vols, const drr_view_t *view,
                float *intensity, float *photon_prob, float *area_density,
                float *solid_angle, void *stream);

/* ---- meshes ------------------------------------------------------------------ */

#define DRR_MESH_MAX_LAYERS 4
#define DRR_MESH_MAX_HITS 128
#define DRR_MESH_RAY_HITS 256

/* One primitive's material (pyrenderdrr/material.py DRRMaterial: density,
 * additive, subtractive, layer, drrMatName). */
typedef struct {
    int32_t material;    /* global material index (< n_materials) */
    float density;       /* g/cm^3 */
    int32_t additive;    /* adds density x chord length to its material */
    int32_t subtractive; /* hides the volumes, and the additive meshes of lower
                            layers, inside it */
    int32_t layer;       /* 0 <= layer < n_layers */
} drr_mesh_prim_t;

/* Triangle meshes of a scene, in world coordinates (mm), ray-cast on the GPU
 * in place of the reference's OpenGL depth peeling (projector.py:1055-1330)
 * and its peel post-processing (peel_postprocess_kernel.cu:15-235).
 * vertices: [n_vertices][3] float; triangles: [n_triangles][3] vertex
 * indices, counter-clockwise seen from outside; triangle_prim: [n_triangles]
 * index into prims.  n_layers: 1..DRR_MESH_MAX_LAYERS (Projector mesh_layers);
 * max_hits: subtractive hits kept per ray and layer, a multiple of 4 in
 * [4, DRR_MESH_MAX_HITS] (max_mesh_hits).  A bounding-volume hierarchy is
 * built on the host and uploaded; the handle owns it. */
int drr_mesh_create(drr_ctx *ctx, const float *vertices, int n_vertices, const int32_t *triangles,
                    const int32_t *triangle_prim, int n_triangles, const drr_mesh_prim_t *prims,
                    int n_prims, int n_layers, int max_hits, drr_mesh **out);
int drr_mesh_destroy(drr_mesh *mesh);

/* drr_project with meshes (project_kernel.cu:363-375, 498-521, 569-579):
 * view->n_volumes may be 0 (meshes only).  source_world: 3 host floats, the
 * source in world coordinates (the rays' origin; alpha is the distance from
 * it along the unit world ray).  source_to_detector_distance (mm, <= 0 if
 * unknown) sets the hit range the reference's peel post-processing keeps
 * (alpha < 2 SDD - 0.001) and the far plane of the additive pass (4 SDD).
 * Outputs as drr_project.  The per-ray mesh terms live in buffers the mesh
 * handle owns, so calls that share a handle must not overlap in time. */
int drr_project_meshes(drr_ctx *ctx, const drr_volume *const *vols, const drr_view_t *view,
                       const drr_mesh *mesh, const float *source_world, float source_to_detector_distance,
                       float *intensity, float *photon_prob, float *area_density, float *solid_angle,
                       void *stream);

/* The per-ray mesh terms drr_project_meshes uses, for inspection:
 * intervals: device float [height][width][max_intervals][2] with
 * max_intervals = n_layers * max_hits / 2, the sorted disjoint ranges
 * (a, b] of alpha hidden by subtractive meshes (any layer inside);
 * counts: device int32 [height][width]; additive: device float
 * [n_layers * n_unique][height][width], per (layer, unique material in
 * ascending global index) the clamped additive area density in mm g/cm^3
 * (0 where the ray's entries and exits do not balance); n_unique (host int,
 * may be NULL) receives the number of unique materials of additive
 * primitives. */
int drr_mesh_rays(drr_ctx *ctx, const drr_mesh *mesh, const drr_view_t *view, const float *source_world,
                  float source_to_detector_distance, float *intervals, int32_t *counts, float *additive,
                  int *n_unique, void *stream);

/* ---- batched forward (one persistent launch, hipGraph) ------------------- */

/* One job of a batch: a view written at `out_offset` floats into the batch
 * outputs.  A view of n_volumes volumes pairs view.vol[v] with the batch's
 * volume volume_index + v (the reference renders every view over all its
 * volumes, projector.py:679-685); views over several volumes run the
 * multi-volume march (priorities, shared label cache) with the same results
 * as drr_project.  The adjoint entry points take single-volume jobs only. */
typedef struct {
    drr_view_t view;           /* n_volumes in [1, 4] */
    int32_t volume_index;      /* first of the view's volumes in the batch's volume list */
    int64_t out_offset;        /* element offset of this view's [H][W] image */
} drr_job_t;

/* Tunables picked by the autotuner (deepdrr_amd/autotune.py). */
typedef struct {
    int32_t tile_w;            /* pixels per work item along u (8 or 16) */
    int32_t tile_h;            /* pixels per work item along v (4 or 8)  */
    int32_t waves_per_block;   /* 1, 2 or 4 */
    int32_t blocks_per_cu;     /* persistent blocks per CU */
} drr_batch_tuning_t;

/* Build a batch plan: uploads the job table, groups ray tiles by volume into
 * 8 per-XCD work queues (with stealing), and captures [queue reset, persistent march, optional neglog epilogue]
 * into a hipGraph.  `neglog` = 1 appends the device neglog of each image
 * (image_utils.py:18-59) writing `images` (float32, same offsets). */
int drr_batch_create(drr_ctx *ctx, const drr_volume *const *vols, int n_vols,
                     const drr_job_t *jobs, int n_jobs, const drr_batch_tuning_t *tuning,
                     float *intensity, float *photon_prob, float *images, int neglog,
                     drr_batch **out);
/* Replay the captured graph on `stream`. */
int drr_batch_launch(drr_batch *batch, void *stream);
/* Kernel-only time (ms) of the last launch's persistent march, from HIP events. */
int drr_batch_last_march_ms(drr_batch *batch, float *ms);
int drr_batch_destroy(drr_batch *batch);

/* One-shot forms of the batch (no graph, job tables built and freed per call;
 * the call returns after the launch completes).  drr_project_jobs writes each
 * job's intensity / photon_prob image at its out_offset (photon_prob may be
 * NULL).  area_density (optional, may be NULL) receives each job's
 * per-material area densities (g/cm^2) as n_materials planes of [height][width]
 * starting at n_materials * out_offset -- what drr_adjoint_jobs can reuse. */
int drr_project_jobs(drr_ctx *ctx, const drr_volume *const *vols, int n_vols, const drr_job_t *jobs,
                     int n_jobs, float *intensity, float *photon_prob, float *area_density, void *stream);

/* drr_project_jobs for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments, checks and
 * results as drr_project_jobs, but the job table, queue state and ray-tile
 * list live in a fixed device array of DRR_NOALLOC_SCRATCH_BYTES that the
 * library holds on every device it runs on until the process ends; nothing
 * is allocated or freed on the device.  A batch whose tables exceed that
 * array (8 bytes per 64-pixel tile plus 152 bytes per job: about 1e6 tiles,
 * or 28 views of 1536^2) is refused with DRR_ERR_INVALID.  All calls on one
 * device share the array, so calls that overlap in time on one device race;
 * each call returns after its launch completes, so calls made one after
 * another (from one thread) never overlap. */
#define DRR_NOALLOC_SCRATCH_BYTES (8u << 20)
int drr_project_jobs_noalloc(drr_ctx *ctx, const drr_volume *const *vols, int n_vols, const drr_job_t *jobs,
                             int n_jobs, float *intensity, float *photon_prob, float *area_density,
                             void *stream);

/* ---- device epilogue --------------------------------------------------------- */

/* Per-image neglog of n images of `npix` pixels at consecutive offsets:
 * x += min(x) + eps; y = -log(x); y = (y - min y) / (max y - min y), or 0
 * for a constant image (image_utils.py:18-59).  In place allowed. */
int drr_neglog(drr_ctx *ctx, const float *in, float *out, int n_images, int64_t npix,
               float epsilon, void *stream);

/* Poisson shot noise with the 3x3 correlation kernel, then clip to
 * [0, 1e31] (analytic_generators.py:10-18).  counter-based RNG (seed, image,
 * pixel).  in/out float32 [n][H][W]. */
int drr_add_noise(drr_ctx *ctx, const float *intensity, const float *photon_prob, float *out,
                  int n_images, int height, int width, float photon_count, uint64_t seed,
                  void *stream);

/* ---- adjoint ----------------------------------------------------------------- */

/* Backward of one view of a single volume (the reference is not
 * differentiable, README.md:47; this is the engine's adjoint of drr_project).
 * With L = sum(grad_intensity * intensity + grad_photon_prob * photon_prob)
 * (either gradient image may be NULL; device float32 [height][width]):
 *   grad_density (device float32 [nz][ny][nx]) += dL/d(density), scattered
 *     atomically with the trilinear weights of every sample;
 *   grad_pose (device float32[6], may be NULL) += dL/d(t_x, t_y, t_z, w_x, w_y, w_z)
 *     for the rigid volume perturbation x -> R(w) (x - c) + c + t applied in
 *     world coordinates, c = pose_center_world, evaluated at w = 0, t = 0.
 * source_world and pose_center_world are 3 host floats.  The per-ray march
 * is recomputed twice (area densities, then gradients); nothing is stored
 * per sample.  Boundary terms of the step count and of the volume entry/exit
 * (piecewise constant in the pose) are not differentiated.  With the scene's
 * attenuate_outside, the air term outside the volume (minAlpha/step and
 * (ray_length - maxAlpha)/step times the air constant, project_kernel.cu:
 * 359-361, 558-560) is held constant as well, although maxAlpha moves with
 * the pose. */
int drr_adjoint(drr_ctx *ctx, const drr_volume *vol, const drr_view_t *view,
                const float *grad_intensity, const float *grad_photon_prob,
                const float *source_world, const float *pose_center_world,
                float *grad_density, float *grad_pose, void *stream);

/* drr_adjoint for many views of one volume in one persistent launch (jobs as
 * in drr_batch_create, all with volume_index 0; the gradient images are read
 * at each job's out_offset).  source_world: 3 floats per job (host).
 * area_density: NULL, or the area densities drr_project_jobs wrote for the
 * same jobs, volume and scene, which saves re-marching the forward pass. */
int drr_adjoint_jobs(drr_ctx *ctx, const drr_volume *vol, const drr_job_t *jobs, int n_jobs,
                     const float *grad_intensity, const float *grad_photon_prob,
                     const float *source_world, const float *pose_center_world,
                     const float *area_density, float *grad_density, float *grad_pose, void *stream);

/* ---- Monte-Carlo scatter ----------------------------------------------------- */

/* Synthetic power-law cross sections and transport settings for the
 * Woodcock-tracking scatter kernel (MC-GPU method; the reference's removed
 * scatter path, projector.py:527-531). */
typedef struct {
    int32_t n_materials;
    int32_t n_bins;
    float energies_keV[DRR_MAX_BINS];
    float cdf[DRR_MAX_BINS];          /* cumulative spectrum pdf, last = 1 */
    /* mu_x[m] (1/mm per g/cm^3) = a_x[m] * (E / 60 keV)^(-k_x) */
    float a_pe[DRR_MAX_MATERIALS], a_compton[DRR_MAX_MATERIALS], a_rayleigh[DRR_MAX_MATERIALS];
    float k_pe, k_compton, k_rayleigh;
    float max_density;                /* unused (the majorants are per 16^3-voxel brick) */
    float source_world[3];            /* point source */
    float world_from_ijk[12];         /* 3x4 row-major */
    float ijk_from_world[12];         /* 3x4 row-major */
    float voxel_size[3];              /* mm */
    float detector_origin[3];         /* world position of pixel (0,0) corner */
    float detector_u[3];              /* world step per pixel along u */
    float detector_v[3];              /* world step per pixel along v */
    float detector_normal[3];         /* unit, pointing away from the source */
    int32_t det_width, det_height;
    float cone_axis[3];               /* unit emission axis */
    float cone_cos;                   /* emit uniformly in cos(theta) >= cone_cos */
    int32_t max_interactions;         /* histories end after this many real collisions */
} drr_mc_params_t;

/* Run histories [first_history, first_history + n_histories) with RANECU
 * per-history streams (seed `seed`; history h always draws the same
 * numbers).  image: device float32 [4][H][W] photon energy (keV) reaching
 * each detector pixel, split as [primary, single Compton, single Rayleigh,
 * multiple] (the column order of MC-GPU's output, test_mcgpu.py:198-235).
 * image2 (optional) receives the per-pixel sums of squared energies (for
 * variance estimates); n_steps (optional, device int64) receives the number
 * of Woodcock flights (each ends in a tentative collision or leaves the
 * volume) and n_brick_steps (optional, device int64) the number of brick
 * majorants they crossed.  All four accumulate.  Delta tracking uses a local
 * majorant per 16x16x16-voxel brick (built per call in stream-ordered memory).
 * The volume is sampled at voxel centres (nearest voxel), IJK box
 * [-0.5, n-0.5]. */
int drr_mc_scatter(drr_ctx *ctx, const drr_volume *vol, const drr_mc_params_t *params,
                   int64_t first_history, int64_t n_histories, uint64_t seed,
                   float *image, float *image2, int64_t *n_steps, int64_t *n_brick_steps, void *stream);

/* ---- timing helper ---------------------------------------------------------- */

/* Milliseconds between two hipEventRecord's around the last drr_project /
 * drr_adjoint / drr_mc_scatter kernel launched by this context. */
int drr_ctx_last_kernel_ms(drr_ctx *ctx, float *ms);

#ifdef __cplusplus
}
#endif
#endif /* DEEPDRR_AMD_H */
