/* mesh_hits.h -- internal interface between oracle/mesh_hits.c and
 * oracle/transcription.c.  TEST INFRASTRUCTURE ONLY. */
#pragma once
#include <stdint.h>

#include "../include/deepdrr_amd.h"

typedef struct {
    const float *vertices;     /* [n_vertices][3], world mm */
    const int32_t *triangles;  /* [n_triangles][3] */
    const int32_t *tri_prim;   /* [n_triangles] */
    int n_triangles;
    const drr_mesh_prim_t *prims;
    int n_prims;
    int n_layers, max_hits;
} ora_mesh;

/* What projectKernel reads per ray (project_kernel.cu:363-375, 569-579). */
typedef struct {
    int8_t layer_valid[DRR_MESH_MAX_LAYERS];                     /* mesh_sub_layer_valid */
    float hit_alpha[DRR_MESH_MAX_LAYERS * DRR_MESH_MAX_HITS];    /* [layer][max_hits] */
    int8_t hit_facing[DRR_MESH_MAX_LAYERS * DRR_MESH_MAX_HITS];
    int n_unique;
    int unique_material[DRR_MAX_MATERIALS];                      /* mesh_unique_materials */
    float additive[DRR_MESH_MAX_LAYERS * DRR_MAX_MATERIALS];     /* [layer][unique] */
    float additive_facing[DRR_MESH_MAX_LAYERS * DRR_MAX_MATERIALS];
} ora_mesh_ray;

int ora_ray_triangle(const float o[3], const float d[3], const float v0[3], const float v1[3],
                     const float v2[3], float *t_out, int *entry);

/* Fill `out` (n_unique and unique_material set by the caller) for the ray
 * from src along the unit direction dir. */
void ora_mesh_ray_terms(const ora_mesh *m, const float src[3], const float dir[3], float sdd,
                        ora_mesh_ray *out);
