// epilogue.h -- device image epilogue shared by the batched path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct ImgSeg {
    int64_t offset;  // first element of the image
    int64_t npix;    // pixels in the image
};

// Enqueue the per-image neglog (reset, min, -log + min/max, normalise) on
// `st`; no allocation, no memset (graph-capturable).  d_keys: 4 uints/image.
int drr_enqueue_neglog(const float *in, float *out, const ImgSeg *d_segs, int n_images,
                       int64_t max_npix, unsigned int *d_keys, float eps, hipStream_t st);
