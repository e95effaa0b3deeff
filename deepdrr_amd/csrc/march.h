// march.h -- per-ray DRR march on CDNA4, shared by the plain, batched and
// adjoint kernels.
//
// Semantics are those of deepdrr's projectKernel
// (deepdrr/projector/project_kernel.cu:195-646): same FP32 operation order,
// IEEE division and sqrt, and no implicit contraction (-ffp-contract=off);
// three multiply-adds per sample are fused explicitly, where nvcc's default
// --fmad=true fuses the reference's expressions: the sample position
// (402-404), the area-density accumulation (543-547) and, in this engine's
// exact-weight trilinear filter, its running sum.  Differences are in *how*
// the work is done, never in the values:
//   * the 8 label corners and 8 density corners of a trilinear cell are read
//     once per cell change from the cell-expanded layout (drr_internal.h) and
//     kept in registers while the ray stays in the cell; the reference keeps
//     the labels (project_kernel.cu:394-431) and re-reads density texels at
//     every step (542);
//   * for a single volume the steps before the volume are skipped with the
//     same sequential FP32 `alpha += step` (552) but without sampling; the
//     label cache is re-established inside a window of at least one voxel
//     before entry, which reproduces the cache state of the reference (the
//     labels a fetch returns depend on the cell alone, see fetch_labels);
//   * a cell whose 8 labels agree sums the weights once instead of per
//     material; a fused multiply-add with a zero fraction leaves an
//     accumulator unchanged, so the result is bit-identical to the
//     per-material sums (451-455, 543-547).
// The fast march (fast_run, DRR_MARCH_FAST) is the one place where values
// differ from the reference's, by rounding only; it is described there.
#pragma once

#include "drr_internal.h"

#define DRR_AIR_DENSITY 0.1129  // double, as project_kernel.cu:6

namespace drr {

__device__ __forceinline__ float add_air(float a) {
    // area_density[AIR] += AIR_DENSITY evaluates in double (project_kernel.cu:526)
    return (float)((double)a + DRR_AIR_DENSITY);
}

__device__ __forceinline__ void ray_dir(const float *W, int udx, int vdx, float &rx, float &ry,
                                        float &rz, float &L) {
    // (float)udx + 0.5 (a double literal) is exact for udx < 2^23.
    const float u = (float)udx + 0.5f;
    const float v = (float)vdx + 0.5f;
    rx = u * W[0] + v * W[1] + W[2];
    ry = u * W[3] + v * W[4] + W[5];
    rz = u * W[6] + v * W[7] + W[8];
    L = sqrtf(rx * rx + ry * ry + rz * rz);
    const float inv = 1.0f / L;
    rx *= inv;
    ry *= inv;
    rz *= inv;
}

// Slab clip of one volume (project_kernel.cu:274-315).  Returns do_trace.
__device__ __forceinline__ bool slab(const float *A, const float *s, const float *lo,
                                     const float *hi, float mrl, float rx, float ry, float rz,
                                     float ri[3], float &mn, float &mx) {
    ri[0] = A[0] * rx + A[1] * ry + A[2] * rz + A[3] * 0.0f;
    ri[1] = A[4] * rx + A[5] * ry + A[6] * rz + A[7] * 0.0f;
    ri[2] = A[8] * rx + A[9] * ry + A[10] * rz + A[11] * 0.0f;
    mn = 0.0f;
    mx = mrl > 0 ? mrl : INFINITY;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (0.0f != ri[a]) {
            const float reci = 1.0f / ri[a];
            const float a0 = (lo[a] - s[a]) * reci;
            const float a1 = (hi[a] - s[a]) * reci;
            mn = fmaxf(mn, fminf(a0, a1));
            mx = fminf(mx, fmaxf(a0, a1));
        } else if (lo[a] > s[a] || s[a] > hi[a]) {
            return false;
        }
    }
    return true;
}

// Sample position along one axis: the reference's s + alpha * r - 1.0f, with
// the multiply-add fused as nvcc fuses it by default (--fmad=true); the
// oracle uses the same fmaf.
__device__ __forceinline__ float sample_pos(float s, float alpha, float r) {
    return __builtin_fmaf(alpha, r, s) - 1.0f;
}

// Volume arrays live in device global memory.  Pointers that reach a kernel
// through a table in memory (the batch's volume list) are generic to the
// compiler, which then emits flat loads (slower, and they also wait on the
// LDS counter); reading through an address-space-1 pointer gives global loads.
typedef const __attribute__((address_space(1))) float gfloat;
typedef const __attribute__((address_space(1))) uint32_t guint;

// Record of cell `cell` (DRR_CELL_WORDS words, drr_internal.h).
__device__ __forceinline__ guint *cell_record(const VolDev &v, uint32_t cell) {
    return (guint *)v.cells + DRR_CELL_WORDS * (size_t)cell;
}

__device__ __forceinline__ int clamp_cell(float b, int n) {
    // cell index in [-1, n-1]; clamp in float (one v_med3_f32) so that huge
    // values stay defined before the conversion
    const float c = __builtin_amdgcn_fmed3f(b, -1.0f, (float)(n - 1));
    return (int)c;
}

// Cells of one volume are < 2^31 (checked at upload), so 32-bit indices suffice.
__device__ __forceinline__ uint32_t cell_index(const VolDev &v, int cx, int cy, int cz) {
    return ((uint32_t)(cz + 1) * (uint32_t)(v.ny + 1) + (uint32_t)(cy + 1)) * (uint32_t)(v.nx + 1) +
           (uint32_t)(cx + 1);
}

// Labels of the 8 corners of the sample at p with base b = floor(p): the
// reference point-samples L[clamp(floor(p + d))] (project_kernel.cu:421-425).
// Every p the march forms is fl(q - 1) for an FP32 q (sample_pos), and for
// such p with |p| < 2^22 the sum p + 1 is exact: for q >= 0.5, q - 1 is exact
// (Sterbenz, and 1 is a multiple of ulp(q) below 2^24), so fl(p + 1) = q; for
// p <= -0.5, |p| - 1 is a multiple of ulp(p).  Hence floor(p + d) = b + d.
// For |p| >= 2^22 the sample is over 2^22 - 16384 voxels outside the volume,
// where b + 1, b + 2 (and b) all clamp to the same edge voxel.  So the labels
// are always L[clamp(b + d)], which is exactly the cell record of
// clamp(b, -1, n-1): l0 = corner labels as nibbles, l1 = common label or ~0.
__device__ __forceinline__ void fetch_labels(const VolDev &v, uint32_t cell, uint32_t &l0, uint32_t &l1) {
    guint *r = cell_record(v, cell);
    l0 = r[8];
    l1 = r[9];
}

__device__ __forceinline__ void weights8(float fx, float fy, float fz, float w[8]) {
    const float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
    w[0] = gx * gy * gz;
    w[1] = fx * gy * gz;
    w[2] = gx * fy * gz;
    w[3] = fx * fy * gz;
    w[4] = gx * gy * fz;
    w[5] = fx * gy * fz;
    w[6] = gx * fy * fz;
    w[7] = fx * fy * fz;
}

__device__ __forceinline__ uint32_t corner_label(uint32_t l0, uint32_t, int d) {
    return (l0 >> (4 * d)) & 0xfu;
}

// seg[m] = sum over corners (dz, dy, dx order) of w[d] where label == m, for a
// cell whose labels differ.
template <int M>
__device__ __forceinline__ void mixed_fractions(uint32_t l0, uint32_t l1, const float w[8], float seg[M]) {
#pragma unroll
    for (int m = 0; m < M; m++) {
        float s = 0.0f;
#pragma unroll
        for (int d = 0; d < 8; d++) s += (corner_label(l0, l1, d) == (uint32_t)m) ? w[d] : 0.0f;
        seg[m] = s;
    }
}

// seg[m] = sum over corners (dz, dy, dx order) of w[d] where label == m.
template <int M>
__device__ __forceinline__ void material_fractions(uint32_t l0, uint32_t l1, const float w[8],
                                                   float seg[M]) {
    const uint32_t m0 = l1;  // the common label, ~0 if the labels differ
    if (m0 != 0xffffffffu) {
        const float ssum = ((((((w[0] + w[1]) + w[2]) + w[3]) + w[4]) + w[5]) + w[6]) + w[7];
#pragma unroll
        for (int m = 0; m < M; m++) seg[m] = ((uint32_t)m == m0) ? ssum : 0.0f;
    } else {
#pragma unroll
        for (int m = 0; m < M; m++) {
            float s = 0.0f;
#pragma unroll
            for (int d = 0; d < 8; d++) s += (corner_label(l0, l1, d) == (uint32_t)m) ? w[d] : 0.0f;
            seg[m] = s;
        }
    }
}

// Density at the sample: corner densities weighted by the same exact FP32
// weights, summed in corner order starting from the first product, each
// further term fused into the running sum.
__device__ __forceinline__ float trilerp(const float4 &d0, const float4 &d1, const float w[8]) {
    float rho = w[0] * d0.x;
    rho = __builtin_fmaf(w[1], d0.y, rho);
    rho = __builtin_fmaf(w[2], d0.z, rho);
    rho = __builtin_fmaf(w[3], d0.w, rho);
    rho = __builtin_fmaf(w[4], d1.x, rho);
    rho = __builtin_fmaf(w[5], d1.y, rho);
    rho = __builtin_fmaf(w[6], d1.z, rho);
    rho = __builtin_fmaf(w[7], d1.w, rho);
    return rho;
}

template <int M>
__device__ __forceinline__ float get_m(const float (&a)[M], int idx) {
    float r = 0.0f;
#pragma unroll
    for (int m = 0; m < M; m++) r = (m == idx) ? a[m] : r;
    return r;
}

template <int M>
__device__ __forceinline__ void set_m(float (&a)[M], int idx, float val) {
#pragma unroll
    for (int m = 0; m < M; m++) a[m] = (m == idx) ? val : a[m];
}

// Cached trilinear cell of the ray: base, 8 labels, 8 densities and whether
// the labels agree (l1 = their common label).  Before the first load the base
// is far from any sample (the fractions p - b test as outside [0, 1)).
struct Cell {
    float bx = -1.0e30f, by = -1.0e30f, bz = -1.0e30f;
    bool have = false;
    bool uni = false;
    uint32_t l0 = 0, l1 = 0;  // corner labels (nibbles); common label or ~0
    float4 d0 = make_float4(0.f, 0.f, 0.f, 0.f), d1 = make_float4(0.f, 0.f, 0.f, 0.f);

    // Cell of base (x, y, z): clamped to [-1, n-1] per axis, its index formed
    // in FP32 (VolDev: exact integers), records read from the biased bases.
    __device__ __forceinline__ void load(const VolDev &vol, float x, float y, float z) {
        const float cx = __builtin_amdgcn_fmed3f(x, -1.0f, vol.cmax[0]);
        const float cy = __builtin_amdgcn_fmed3f(y, -1.0f, vol.cmax[1]);
        const float cz = __builtin_amdgcn_fmed3f(z, -1.0f, vol.cmax[2]);
        const uint32_t ib = cell_index_biased(vol, cx, cy, cz);  // index + 0x4B000000
        gfloat *rp = (gfloat *)(vol.cells_b + (uint64_t)ib * (4u * DRR_CELL_WORDS));
        d0 = make_float4(rp[0], rp[1], rp[2], rp[3]);
        d1 = make_float4(rp[4], rp[5], rp[6], rp[7]);
        l0 = ((guint *)rp)[8];
        l1 = ((guint *)rp)[9];
        uni = l1 != 0xffffffffu;
        bx = x;
        by = y;
        bz = z;
        have = true;
    }
    __device__ __forceinline__ void load(const VolDev &vol, float, float, float, float x, float y, float z) {
        load(vol, x, y, z);
    }
};

// Running area density of the current uniform material: ad[cm] lives in
// `cur` across consecutive uniform cells of one material and is written back
// only when the material changes (cm < 0: none).
struct Accum {
    float cur = 0.0f;
    int cm = -1;
};

// One in-volume sample of the single-volume march at position p
// (project_kernel.cu:394-547) with weight wt (0.5 at the ends of the ray, else 1).
template <int M>
__device__ __forceinline__ void sample_v1(const VolDev &vol, Cell &c, float (&ad)[M], Accum &a, float px,
                                          float py, float pz, float wt) {
    // floor(p) is the cached base exactly when every fraction p - b lies in
    // [0, 1): p - b is exact there, and rounding is monotonic.  As unsigned
    // bits, [0, 1) is [0, 0x3f800000); -0, negatives and NaN fall outside.
    float fx = px - c.bx, fy = py - c.by, fz = pz - c.bz;
    const uint32_t fmax = max(max(__float_as_uint(fx), __float_as_uint(fy)), __float_as_uint(fz));
    if (fmax >= 0x3f800000u) {
        const float bx = floorf(px), by = floorf(py), bz = floorf(pz);
        c.load(vol, bx, by, bz);
        fx = px - bx;
        fy = py - by;
        fz = pz - bz;
        const int ncm = (int)c.l1;  // the common material, -1 if the labels differ
        if (ncm != a.cm) {
            if (a.cm >= 0) set_m<M>(ad, a.cm, a.cur);
            a.cm = ncm;
            a.cur = get_m<M>(ad, ncm);
        }
    }
    float w[8];
    weights8(fx, fy, fz, w);
    const float rho = trilerp(c.d0, c.d1, w);
    const float x = wt * rho;
    const float ssum = ((((((w[0] + w[1]) + w[2]) + w[3]) + w[4]) + w[5]) + w[6]) + w[7];
    // uniform cell: seg[m0] = ssum, all other fractions 0
    a.cur = c.uni ? __builtin_fmaf(x, ssum, a.cur) : a.cur;
    if (!c.uni) {
        // Labels differ: per material present among the 8 corners (a zero
        // nibble of l0 ^ m * 0x11111111, by the borrow test), its fraction in
        // corner order; an absent material's fraction is 0 and leaves ad[m]
        // unchanged (fma(x, 0, a) = a for a >= 0), so it is skipped.
#pragma unroll
        for (int m = 0; m < M; m++) {
            const uint32_t xm = c.l0 ^ (0x11111111u * (uint32_t)m);
            if (((xm - 0x11111111u) & ~xm & 0x88888888u) != 0u) {
                float sm = 0.0f;
#pragma unroll
                for (int d = 0; d < 8; d++) sm += (corner_label(c.l0, c.l1, d) == (uint32_t)m) ? w[d] : 0.0f;
                ad[m] = __builtin_fmaf(x, sm, ad[m]);
            }
        }
    }
}

// The fast march (DRR_MARCH_FAST): the samples t in [t, tlim), all of weight 1,
// with the same FP32 alpha sequence as sample_v1 and, in cells whose labels
// agree, the same sequential FP32 additions into the area densities, but with
// each sample's density formed far more cheaply than by the exact filter.
// Along the ray the trilinear density of a cell is a cubic in the distance
// u = alpha - a0 from the sample where the ray entered the cell; on entering a
// cell its four coefficients are formed from the 8 corner densities, the entry
// fractions and the ray direction, with the distance to the exit face, and each
// sample then costs three multiply-adds and the loop's bookkeeping.  The entry
// sample uses the exact march's position and floor; the samples after it are
// placed by the exact alpha sequence (u is exact) on the exact line instead of
// at their rounded FP32 positions, and a sample within rounding of a face may
// take either cell (the density is continuous there).  A cell with two labels
// splits the density by the second label's fraction, also a cubic (fast_run);
// a cell with more takes the exact filter's formulas.  Each sample's value thus
// differs from the exact march's by a few roundings, with both signs, and the
// sums stay within about 1e-7 relative of the exact march's; nothing depends
// on the launch or on other lanes, so results are the same from run to run and
// from kernel to kernel.

// Coefficients of the trilinear interpolant of corner values (v0 = corners
// 0..3, v1 = 4..7, corner d = dx + 2 dy + 4 dz) along the ray p(u) = p0 + u r
// from fractions f = p0 - b: value(u) = ((k3 u + k2) u + k1) u + k0.
__device__ __forceinline__ void cell_cubic(const float4 &v0, const float4 &v1, float fx, float fy, float fz,
                                           float rx, float ry, float rz, float &k0, float &k1, float &k2,
                                           float &k3) {
    // x: e_yz(u) = A_yz + B_yz u, the lerp of corners (0yz, 1yz)
    const float X00 = v0.y - v0.x, X10 = v0.w - v0.z, X01 = v1.y - v1.x, X11 = v1.w - v1.z;
    const float A00 = __builtin_fmaf(fx, X00, v0.x), A10 = __builtin_fmaf(fx, X10, v0.z);
    const float A01 = __builtin_fmaf(fx, X01, v1.x), A11 = __builtin_fmaf(fx, X11, v1.z);
    const float B00 = rx * X00, B10 = rx * X10, B01 = rx * X01, B11 = rx * X11;
    // y: g_z(u) = G_z0 + G_z1 u + G_z2 u^2
    const float P0 = A10 - A00, Q0 = B10 - B00, P1 = A11 - A01, Q1 = B11 - B01;
    const float G00 = __builtin_fmaf(fy, P0, A00), G01 = __builtin_fmaf(fy, Q0, __builtin_fmaf(ry, P0, B00));
    const float G02 = ry * Q0;
    const float G10 = __builtin_fmaf(fy, P1, A01), G11 = __builtin_fmaf(fy, Q1, __builtin_fmaf(ry, P1, B01));
    const float G12 = ry * Q1;
    // z: the cubic
    const float D0 = G10 - G00, D1 = G11 - G01, D2 = G12 - G02;
    k0 = __builtin_fmaf(fz, D0, G00);
    k1 = __builtin_fmaf(fz, D1, __builtin_fmaf(rz, D0, G01));
    k2 = __builtin_fmaf(fz, D2, __builtin_fmaf(rz, D1, G02));
    k3 = rz * D2;
}

struct FastCell {
    // density = ((c3 u + c2) u + c1) u + c0 (fast_run keeps the second label's
    // corner indicators here instead in a two-label cell, zeros in a cell with more)
    float c0 = 0.f, c1 = 0.f, c2 = 0.f, c3 = 0.f;
    float a0 = 0.f;    // alpha of the entry sample (u = alpha - a0)
    float nex = 0.f;   // distance to the exit face, in steps

    // Enter cell c at alpha with fractions (fx, fy, fz) (those of the exact
    // march) and the ray direction r in voxels per mm.  Per axis, the exit
    // face is reached after fma(-f, j, h) steps with j = 1 / (r step) and
    // h = j going up, 0 going down (j = 0, h = +inf where r = 0).
    __device__ __forceinline__ void enter(const Cell &c, float alpha, float fx, float fy, float fz, float rx,
                                          float ry, float rz, float jx, float jy, float jz, float hx, float hy,
                                          float hz) {
        a0 = alpha;
        cell_cubic(c.d0, c.d1, fx, fy, fz, rx, ry, rz, c0, c1, c2, c3);
        nex = fminf(__builtin_fmaf(-fx, jx, hx), fminf(__builtin_fmaf(-fy, jy, hy), __builtin_fmaf(-fz, jz, hz)));
    }
};

// Bit 3 of nibble d is set where nibble d of x is nonzero (exact per nibble:
// the low three bits plus 7 carry into bit 3 and never into the next nibble).
__device__ __forceinline__ uint32_t nz_nibbles(uint32_t x) {
    return (((x & 0x77777777u) + 0x77777777u) | x) & 0x88888888u;
}

// The samples of the entered cells (fast_run): n for this lane, the density
// cubic in fc (tw: a two-label cell, whose corner data fc and e hold instead)
// and the two-label formulas where tw (TWO: some lane of the wave has a
// two-label cell; the two forms are separate loops).
template <bool TWO>
__device__ __forceinline__ void fast_cell(const Cell &c, const FastCell &fc, Accum &a, float &sq, float e0,
                                          float e1, float e2, float e3, bool tw, int n, float sx, float sy,
                                          float sz, float rix, float riy, float riz, float step, float &alpha) {
    float u = 0.0f;
    for (int k = 0;; ++k) {
        const bool in = k < n;
        if (__ballot(in) == 0) break;
        float rho = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(fc.c3, u, fc.c2), u, fc.c1), u, fc.c0);
        const float g = in ? 1.0f : 0.0f;
        if (TWO) {
            // two labels: density and q-fraction at the sample's FP32 position
            // (that of the exact filter), each by nested lerps
            const float fx = __builtin_fmaf(alpha, rix, sx) - (c.bx + 1.0f);
            const float fy = __builtin_fmaf(alpha, riy, sy) - (c.by + 1.0f);
            const float fz = __builtin_fmaf(alpha, riz, sz) - (c.bz + 1.0f);
            const float a00 = __builtin_fmaf(fx, c.d0.y - c.d0.x, c.d0.x);
            const float a10 = __builtin_fmaf(fx, c.d0.w - c.d0.z, c.d0.z);
            const float a01 = __builtin_fmaf(fx, c.d1.y - c.d1.x, c.d1.x);
            const float a11 = __builtin_fmaf(fx, c.d1.w - c.d1.z, c.d1.z);
            const float b0 = __builtin_fmaf(fy, a10 - a00, a00), b1 = __builtin_fmaf(fy, a11 - a01, a01);
            const float x = __builtin_fmaf(fz, b1 - b0, b0);
            const float i00 = __builtin_fmaf(fx, e0, fc.c0), i10 = __builtin_fmaf(fx, e1, fc.c1);
            const float i01 = __builtin_fmaf(fx, e2, fc.c2), i11 = __builtin_fmaf(fx, e3, fc.c3);
            const float j0 = __builtin_fmaf(fy, i10 - i00, i00), j1 = __builtin_fmaf(fy, i11 - i01, i01);
            const float fq = __builtin_fmaf(fz, j1 - j0, j0);
            const float xq = x * fq;
            sq = __builtin_fmaf(g, tw ? xq : 0.0f, sq);
            rho = tw ? x - xq : rho;
        }
        a.cur = __builtin_fmaf(g, rho, a.cur);
        alpha = __builtin_fmaf(g, step, alpha);
        u = alpha - fc.a0;
    }
}

// The samples of a cell with three or more labels (ex), after fast_cell has
// taken the cell's samples for the other lanes: the exact filter's formulas at
// the same FP32 alpha sequence, re-walked from the entry sample a0, added into
// ad[] sample by sample (the lane has no running sum in such a cell).
template <int M>
__device__ __forceinline__ void fast_exact(const Cell &c, float (&ad)[M], bool ex, int n, float a0, float sx,
                                           float sy, float sz, float rix, float riy, float riz, float step) {
    float al = a0;
    for (int k = 0;; ++k) {
        const bool in = (k < n) & ex;
        if (__ballot(in) == 0) break;
        if (in) {
            const float fx = __builtin_fmaf(al, rix, sx) - (c.bx + 1.0f);
            const float fy = __builtin_fmaf(al, riy, sy) - (c.by + 1.0f);
            const float fz = __builtin_fmaf(al, riz, sz) - (c.bz + 1.0f);
            float w[8];
            weights8(fx, fy, fz, w);
            const float x = trilerp(c.d0, c.d1, w);
#pragma unroll
            for (int m = 0; m < M; m++) {
                const uint32_t xm = c.l0 ^ (0x11111111u * (uint32_t)m);
                if (((xm - 0x11111111u) & ~xm & 0x88888888u) != 0u) {
                    float sm = 0.0f;
#pragma unroll
                    for (int d = 0; d < 8; d++) sm += (corner_label(c.l0, c.l1, d) == (uint32_t)m) ? w[d] : 0.0f;
                    ad[m] = __builtin_fmaf(x, sm, ad[m]);
                }
            }
        }
        al += step;
    }
}

template <int M>
__device__ __forceinline__ void fast_run(const VolDev &vol, Cell &c, FastCell &fc, float (&ad)[M], Accum &a,
                                         float sx, float sy, float sz, float rix, float riy, float riz,
                                         float step, float &alpha, int &t, int tlim, float amax) {
    // Cell by cell: every lane still marching enters the cell of its current
    // sample together (one execution of the long cell-entry block for the whole
    // wave), then the wave runs the samples of those cells until every lane has
    // taken its own.  Lanes whose cells end early idle for the rest of that inner
    // loop, which costs less than running the entry block once for each group
    // of lanes that leave together.  A cell's sample count n is fixed on entry:
    // the samples at u = 0, step, 2 step, ... (for a far source, at the FP32
    // increment) before the exit distance (at least the entry sample, at
    // most the samples left before tlim); a sample within the FP32 alpha
    // sequence's drift of the face may so fall to either cell, where the
    // density is continuous.  The inner loop's test is a wave vote and
    // its per-lane choices are 0/1 factors: per-lane branches and loop exits cost
    // scalar instructions (exec-mask bookkeeping), enough to become the limit in
    // a loop this short.  With g = 1, fma(g, rho, cur) = cur + rho and
    // fma(g, step, alpha) = alpha + step exactly; with g = 0 both leave the
    // value unchanged.
    //
    // Materials: a cell whose labels agree adds each sample's density to the
    // running sum of its material (Accum).  A cell with exactly two labels
    // (nearly every cell on a material boundary) forms each sample's density
    // and the fraction f_q of its second label q at the sample's FP32 position,
    // as the exact filter does (across a boundary the density changes by up to
    // several g/cm^3 per voxel, so that position's rounding shows at 1e-5 in
    // the sums, unlike in a cell of one material), each by nested lerps:
    // rho (1 - f_q) goes to the running sum of the first material and rho f_q
    // to a second running sum sq of ad[q], written back when q or the first
    // material changes and when the run ends; so every term is still added into
    // its material's area density one sample at a time, as the exact filter
    // adds it.  A cell with three or more labels takes the exact filter's
    // formulas, its density cubic set to zero.  In a wave with no two-label
    // cell the two-label formulas are skipped; a lane in a one-label cell
    // selects its cubic's value and adds zero to sq, so its results do not
    // depend on the other lanes.
    if (__ballot(t < tlim) == 0) return;
    // exit distances in steps (reciprocals to 1 ulp are plenty there), kept in
    // registers (an empty asm stops the compiler from forming them again on
    // every cell entry)
    const float istep = 1.0f / step;
    float jx = rix == 0.0f ? 0.0f : __builtin_amdgcn_rcpf(rix) * istep;
    float jy = riy == 0.0f ? 0.0f : __builtin_amdgcn_rcpf(riy) * istep;
    float jz = riz == 0.0f ? 0.0f : __builtin_amdgcn_rcpf(riz) * istep;
    float hx = rix == 0.0f ? INFINITY : (rix < 0.0f ? 0.0f : jx);
    float hy = riy == 0.0f ? INFINITY : (riy < 0.0f ? 0.0f : jy);
    float hz = riz == 0.0f ? INFINITY : (riz < 0.0f ? 0.0f : jz);
    // Each FP32 `alpha += step` moves alpha by step rounded to alpha's ulp, off
    // step by up to half an ulp.  Below 2^14 steps (about 1.6 m at 0.1 mm) that
    // is under 2^-11 of a step and the samples a cell's count assumes stay
    // within rounding of its face.  Beyond (a far source: at 4 km alpha moves a
    // quarter step more, or not at all) the exit distances are scaled by step
    // over the larger of the increments at the ray's current alpha and at its
    // largest, those of the two binades a far ray spans, so a cell's count ends
    // at its exit face (a ray that also crosses lower binades, starting near its
    // source, has their increments within 3/4 of an ulp of amax of these); where
    // alpha stops moving, every entry takes one sample.
    if (amax > 16384.0f * step) {
        const float s_lo = (alpha + step) - alpha, s_hi = (amax + step) - amax;
        const float fac = (s_lo > 0.0f && s_hi > 0.0f) ? step / fmaxf(s_lo, s_hi) : 0.0f;
        jx *= fac;
        jy *= fac;
        jz *= fac;
        hx = fac > 0.0f ? hx * fac : 0.0f;
        hy = fac > 0.0f ? hy * fac : 0.0f;
        hz = fac > 0.0f ? hz * fac : 0.0f;
    }
    asm volatile("" : "+v"(jx), "+v"(jy), "+v"(jz), "+v"(hx), "+v"(hy), "+v"(hz));
    float sq = 0.0f;  // ad[q] while q (the second material of two-label cells) is >= 0
    int q = -1;
    // two labels: the q-indicators' x-differences (read only where tw, so not
    // cleared for other cells)
    float e0 = 0.f, e1 = 0.f, e2 = 0.f, e3 = 0.f;
    do {
        int n = 0;        // samples of the entered cell
        bool ex = false;  // three or more labels: the exact filter
        bool tw = false;  // two labels
        if (t < tlim) {
            const float px = sample_pos(sx, alpha, rix), py = sample_pos(sy, alpha, riy),
                        pz = sample_pos(sz, alpha, riz);
            const float bx = floorf(px), by = floorf(py), bz = floorf(pz);
            c.load(vol, bx, by, bz);
            const float fx = px - bx, fy = py - by, fz = pz - bz;
            fc.enter(c, alpha, fx, fy, fz, rix, riy, riz, jx, jy, jz, hx, hy, hz);
            int ncm = (int)c.l1;  // the first material (the common one in a one-label cell)
            int qn = -1;          // the second material of a two-label cell
            if (!c.uni) {
                const uint32_t l0 = c.l0;
                const uint32_t la = l0 & 0xfu;
                const uint32_t xa = l0 ^ (la * 0x11111111u);  // nonzero nibbles: corners not la
                const uint32_t lb = (l0 >> (__builtin_ctz(xa) & ~3u)) & 0xfu;
                const uint32_t xb = l0 ^ (lb * 0x11111111u);
                const uint32_t na = nz_nibbles(xa), nb = nz_nibbles(xb);
                if ((na & nb) == 0u) {
                    // the first material keeps the running sum when it already has it
                    const bool sw = (int)lb == a.cm;
                    ncm = (int)(sw ? lb : la);
                    qn = (int)(sw ? la : lb);
                    const uint32_t iq = sw ? na : nb;  // bit set: corner not q
                    // q-indicators of the corners: bases of the x-lerps in fc.c0..c3,
                    // their differences in e0..e3
                    fc.c0 = (iq & 0x8u) ? 0.f : 1.f;
                    fc.c1 = (iq & 0x800u) ? 0.f : 1.f;
                    fc.c2 = (iq & 0x80000u) ? 0.f : 1.f;
                    fc.c3 = (iq & 0x8000000u) ? 0.f : 1.f;
                    e0 = ((iq & 0x80u) ? 0.f : 1.f) - fc.c0;
                    e1 = ((iq & 0x8000u) ? 0.f : 1.f) - fc.c1;
                    e2 = ((iq & 0x800000u) ? 0.f : 1.f) - fc.c2;
                    e3 = ((iq & 0x80000000u) ? 0.f : 1.f) - fc.c3;
                    tw = true;
                } else {
                    ex = true;
                    fc.c0 = fc.c1 = fc.c2 = fc.c3 = 0.0f;
                }
            }
            if (q >= 0 && (ncm != a.cm || (qn >= 0 && qn != q))) {
                set_m<M>(ad, q, sq);
                q = -1;
            }
            if (ncm != a.cm) {
                if (a.cm >= 0) set_m<M>(ad, a.cm, a.cur);
                a.cm = ncm;
                a.cur = get_m<M>(ad, ncm);
            }
            if (qn >= 0 && q < 0) {
                q = qn;
                sq = get_m<M>(ad, qn);
            }
            // the samples before the exit face, clamped to [1, tlim - t] after
            // the conversion (which saturates infinities and takes NaN to 0), so
            // every entry takes at least one sample
            n = min(max((int)ceilf(fc.nex), 1), tlim - t);
        }
        if (__ballot(tw) != 0)
            fast_cell<true>(c, fc, a, sq, e0, e1, e2, e3, tw, n, sx, sy, sz, rix, riy, riz, step, alpha);
        else
            fast_cell<false>(c, fc, a, sq, e0, e1, e2, e3, tw, n, sx, sy, sz, rix, riy, riz, step, alpha);
        if (__ballot(ex) != 0) fast_exact<M>(c, ad, ex, n, fc.a0, sx, sy, sz, rix, riy, riz, step);
        t += n;
    } while (__ballot(t < tlim) != 0);
    if (q >= 0) set_m<M>(ad, q, sq);
    // sample_v1 adds a mixed cell's terms into ad[] directly and expects no
    // running sum there: write a two-label cell's first material back
    if (!c.uni && a.cm >= 0) {
        set_m<M>(ad, a.cm, a.cur);
        a.cm = -1;
        a.cur = 0.0f;
    }
}

// Occupancy target of the march kernels: up to 5 materials the march fits in
// 64 VGPRs (8 waves per SIMD) without spills in its sample loops; more
// materials get more registers.
template <int M>
struct MarchOcc {
    static constexpr int waves = M <= 5 ? 8 : 5;
};

// Per-ray state of the single-volume march, filled by march_setup_v1.
struct RayV1 {
    float rx, ry, rz, L;
    float ri[3];
    float mnv, mxv;
    float minA, maxA;
    int N;
    bool trace;
};

// `while (t < N && alpha < lim) { alpha += step; ++t; }` with the same FP32
// additions, tested once per 8 steps: alpha only grows, so the 8 tests of a
// block pass exactly when the one before its last addition does.
__device__ __forceinline__ void skip_steps(int N, float lim, float step, float &alpha, int &t) {
    while (t + 8 <= N) {
        float a = alpha;
#pragma unroll
        for (int k = 0; k < 7; k++) a += step;
        if (!(a < lim)) break;
        alpha = a + step;
        t += 8;
    }
    while (t < N && alpha < lim) {
        alpha += step;
        ++t;
    }
}

__device__ __forceinline__ void march_setup_v1(const Geo1 &g, int udx, int vdx, float step, RayV1 &r) {
    ray_dir(g.W, udx, vdx, r.rx, r.ry, r.rz, r.L);
    r.minA = r.L;
    r.maxA = 0.0f;
    r.trace = false;
    r.ri[0] = r.ri[1] = r.ri[2] = 0.0f;
    r.mnv = 0.0f;
    r.mxv = 0.0f;
    if (g.en) {
        r.trace = slab(g.A, g.s, g.lo, g.hi, g.mrl, r.rx, r.ry, r.rz, r.ri, r.mnv, r.mxv);
        if (r.trace) {
            r.minA = fminf(r.minA, r.mnv);
            r.maxA = fmaxf(r.maxA, r.mxv);
        }
    }
    r.N = (int)ceilf((r.maxA - r.minA) / step);
}

// Single-volume march (project_kernel.cu:334-584 with NUM_VOLUMES == 1).
// On return ad[] holds the final area densities in g/cm^2.
template <int M, bool ATTEN, bool FAST = false>
__device__ __forceinline__ void march_v1(const Geo1 &g, const VolDev &vol, int udx, int vdx,
                                         float step, int air, float (&ad)[M]) {
    RayV1 r;
    march_setup_v1(g, udx, vdx, step, r);
#pragma unroll
    for (int m = 0; m < M; m++) ad[m] = 0.0f;
    float acc = 0.0f;  // air accumulator while no volume is sampled
    if (ATTEN) acc = (float)((double)acc + (double)(r.minA / step) * DRR_AIR_DENSITY);

    const int N = r.N;
    float alpha = r.minA;
    int t = 0;
    if (r.trace) {
        const float sx = g.s[0], sy = g.s[1], sz = g.s[2];
        const float rix = r.ri[0], riy = r.ri[1], riz = r.ri[2];
        // Phase 1: steps well before the volume: only the alpha sequence.
        const float rdom = fmaxf(fabsf(rix), fmaxf(fabsf(riy), fabsf(riz)));
        const float lim = r.mnv - (1.0f / rdom + 2.0f * step);
        if (ATTEN) {
            while (t < N && alpha < lim) {
                acc = add_air(acc);
                alpha += step;
                ++t;
            }
        } else {
            skip_steps(N, lim, step, alpha, t);
        }
        // Phase 2: approach window, label cache tracked as the reference does.
        Cell c;
        while (t < N && alpha < r.mnv) {
            const float px = sample_pos(sx, alpha, rix);
            const float py = sample_pos(sy, alpha, riy);
            const float pz = sample_pos(sz, alpha, riz);
            const float bx = floorf(px), by = floorf(py), bz = floorf(pz);
            if (!c.have || bx != c.bx || by != c.by || bz != c.bz) c.load(vol, px, py, pz, bx, by, bz);
            if (ATTEN) acc = add_air(acc);
            alpha += step;
            ++t;
        }
        if (ATTEN) set_m<M>(ad, air, acc);
        // Phase 3: inside [minAlpha_vol, maxAlpha_vol]: sample and accumulate.
        // While the ray is in a cell whose 8 labels agree (material m0), only
        // ad[m0] changes; it is kept in `cur` and written back on leaving, so
        // every FP32 addition is the one the reference performs.  The end
        // samples t == 0 and t == N-1 carry weight 0.5 (project_kernel.cu:
        // 541), the others 1.
        Accum acc3;
        if (c.have && c.uni) {
            acc3.cm = (int)c.l1;
            acc3.cur = get_m<M>(ad, acc3.cm);
        }
        // The loop runs while t < tlim with no test on alpha: tlim is at most
        // the last weight-1 sample (N - 2) and at most the samples whose alpha
        // is certainly <= maxAlpha_vol.  Each FP32 `alpha += step` adds at
        // most step + 2^-24 (|mxv| + step) while alpha <= mxv, so with
        // inc = step + 2^-23 (|mxv| + step) the next K = floor((mxv - alpha)
        // / inc) - 2 samples stay inside (the -2 covers the FP32 quotient and
        // difference).  Past tlim the loop re-enters one sample at a time with
        // the full test; the last sample (t == N-1) re-enters with weight 0.5.
        float wt = (0 == t) ? 0.5f : 1.0f;
        const float inc = step + 1.1920929e-7f * (fabsf(r.mxv) + step);
        const float kf = (r.mxv - alpha) / inc - 2.0f;
        // (clamped in float to a range the conversion holds, then to N - 1 in
        // integers, where (float)(N - 1 - t) could round up)
        int tlim = min(t + (int)__builtin_amdgcn_fmed3f(kf, 0.0f, 16777216.0f), N - 1);
        for (;;) {
            if (FAST) {
                // a sample of weight 0.5 (the first) takes the exact filter
                if (wt != 1.0f && t < tlim) {
                    sample_v1<M>(vol, c, ad, acc3, sample_pos(sx, alpha, rix), sample_pos(sy, alpha, riy),
                                 sample_pos(sz, alpha, riz), wt);
                    wt = 1.0f;
                    alpha += step;
                    ++t;
                }
                FastCell fc;  // no cell entered: the first sample enters one
                fast_run<M>(vol, c, fc, ad, acc3, sx, sy, sz, rix, riy, riz, step, alpha, t, tlim, r.mxv);
            } else {
                while (t < tlim) {
                    sample_v1<M>(vol, c, ad, acc3, sample_pos(sx, alpha, rix), sample_pos(sy, alpha, riy),
                                 sample_pos(sz, alpha, riz), wt);
                    wt = 1.0f;
                    alpha += step;
                    ++t;
                }
            }
            // a fresh comparison after the loop (an empty asm hides alpha's
            // in-loop history), so the loop does not carry one
            asm volatile("" : "+v"(alpha));
            if (!(alpha <= r.mxv) || t >= N) break;
            if (t == N - 1) wt = 0.5f;
            tlim = t + 1;
        }
        if (acc3.cm >= 0) set_m<M>(ad, acc3.cm, acc3.cur);
        if (ATTEN) acc = get_m<M>(ad, air);
    }
    // Phase 4: the rest of the steps sample nothing.
    if (ATTEN) {
        for (; t < N; ++t) acc = add_air(acc);
        acc = (float)((double)acc + (double)((r.L - r.maxA) / step) * DRR_AIR_DENSITY);
        set_m<M>(ad, air, acc);
    }
#pragma unroll
    for (int m = 0; m < M; m++) ad[m] *= step;
#pragma unroll
    for (int m = 0; m < M; m++) ad[m] /= 10.0f;
}

// Mesh terms of a ray for march_multi: `inside(alpha)` says whether a
// subtractive mesh hides the sample at alpha (called with increasing alpha,
// project_kernel.cu:498-521) and `add(ad)` adds the additive mesh area
// densities after the step scaling (569-579).  NoMesh is the scene without
// meshes.
struct NoMesh {
    __device__ __forceinline__ bool inside(float) { return false; }
    template <int M>
    __device__ __forceinline__ void add(float (&)[M]) {}
};

// General march for 1..V volumes with priorities and the label cache shared
// across volumes (project_kernel.cu:394-553), literal step by step.
template <int V, int M, typename Mesh = NoMesh>
__device__ __forceinline__ void march_multi(const ViewDev &vd, int udx, int vdx, float step,
                                            int atten, int air, float (&ad)[M], Mesh mesh = Mesh()) {
    float rx, ry, rz, L;
    ray_dir(vd.W, udx, vdx, rx, ry, rz, L);
    float minA = L, maxA = 0.0f;
#pragma unroll
    for (int m = 0; m < M; m++) ad[m] = 0.0f;
    constexpr int VA = V > 0 ? V : 1;  // array extent (V = 0: meshes only)
    float ri[VA][3], mnv[VA], mxv[VA];
    bool trace[VA];
#pragma unroll
    for (int i = 0; i < V; i++) {
        ri[i][0] = ri[i][1] = ri[i][2] = 0.0f;
        mnv[i] = mxv[i] = 0.0f;
        trace[i] = false;
        if (!vd.en[i]) continue;
        trace[i] = slab(vd.A[i], vd.s[i], vd.lo[i], vd.hi[i], vd.max_ray_length, rx, ry, rz, ri[i],
                        mnv[i], mxv[i]);
        if (trace[i]) {
            minA = fminf(minA, mnv[i]);
            maxA = fmaxf(maxA, mxv[i]);
        }
    }
    const int N = (int)ceilf((maxA - minA) / step);
    float acc_air = 0.0f;
    if (atten) acc_air = (float)((double)acc_air + (double)(minA / step) * DRR_AIR_DENSITY);
    if (atten) set_m<M>(ad, air, acc_air);

    float pbx = 0.0f, pby = 0.0f, pbz = 0.0f;
    bool have = false;
    uint32_t l0 = 0, l1 = 0;
    float alpha = minA;
    for (int t = 0; t < N; t++) {
        float p[VA][3];
        float seg[VA][M];
#pragma unroll
        for (int i = 0; i < V; i++) {
#pragma unroll
            for (int m = 0; m < M; m++) seg[i][m] = 0.0f;
            p[i][0] = p[i][1] = p[i][2] = 0.0f;
            if (!vd.en[i]) continue;
            const VolDev &vol = vd.vol[i];
            p[i][0] = sample_pos(vd.s[i][0], alpha, ri[i][0]);
            p[i][1] = sample_pos(vd.s[i][1], alpha, ri[i][1]);
            p[i][2] = sample_pos(vd.s[i][2], alpha, ri[i][2]);
            const float bx = floorf(p[i][0]), by = floorf(p[i][1]), bz = floorf(p[i][2]);
            if (!have || bx != pbx || by != pby || bz != pbz) {
                const uint32_t cell = cell_index(vol, clamp_cell(bx, vol.nx), clamp_cell(by, vol.ny),
                                                clamp_cell(bz, vol.nz));
                fetch_labels(vol, cell, l0, l1);
                pbx = bx;
                pby = by;
                pbz = bz;
                have = true;
            }
            float w[8];
            weights8(p[i][0] - bx, p[i][1] - by, p[i][2] - bz, w);
            material_fractions<M>(l0, l1, w, seg[i]);
        }
        int cp = V, n_at = 0;
#pragma unroll
        for (int i = 0; i < V; i++) {
            if (!trace[i] || !vd.en[i]) continue;
            if ((alpha < mnv[i]) || (alpha > mxv[i])) continue;
            float any = 0.0f;
#pragma unroll
            for (int m = 0; m < M; m++) any += seg[i][m];
            // any > 0 whenever a sum prefix is > 0 (all terms are >= 0)
            if (0.0f == any) continue;
            if (vd.prio[i] < cp) {
                cp = vd.prio[i];
                n_at = 1;
            } else if (vd.prio[i] == cp) {
                n_at += 1;
            }
        }
        if (mesh.inside(alpha)) {
            // the volumes (and the outside air) are hidden inside a mesh
        } else if (0 == n_at) {
            if (atten) set_m<M>(ad, air, add_air(get_m<M>(ad, air)));
        } else {
            float weight = 1.0f / ((float)n_at);
            weight *= (0 == t || N - 1 == t) ? 0.5f : 1.0f;
#pragma unroll
            for (int i = 0; i < V; i++) {
                if (trace[i] && vd.prio[i] == cp && vd.en[i] == 1) {
                    const VolDev &vol = vd.vol[i];
                    const float bx = floorf(p[i][0]), by = floorf(p[i][1]), bz = floorf(p[i][2]);
                    const uint32_t cell = cell_index(vol, clamp_cell(bx, vol.nx), clamp_cell(by, vol.ny),
                                                    clamp_cell(bz, vol.nz));
                    gfloat *rp = (gfloat *)cell_record(vol, cell);
                    const float4 d0 = make_float4(rp[0], rp[1], rp[2], rp[3]);
                    const float4 d1 = make_float4(rp[4], rp[5], rp[6], rp[7]);
                    float w[8];
                    weights8(p[i][0] - bx, p[i][1] - by, p[i][2] - bz, w);
                    const float rho = trilerp(d0, d1, w);
#pragma unroll
                    for (int m = 0; m < M; m++) ad[m] = __builtin_fmaf(weight * rho, seg[i][m], ad[m]);
                }
            }
        }
        alpha += step;
    }
    if (atten) {
        float a = get_m<M>(ad, air);
        a = (float)((double)a + (double)((L - maxA) / step) * DRR_AIR_DENSITY);
        set_m<M>(ad, air, a);
    }
#pragma unroll
    for (int m = 0; m < M; m++) ad[m] *= step;
    mesh.add(ad);
#pragma unroll
    for (int m = 0; m < M; m++) ad[m] /= 10.0f;
}

// Spectral Beer-Lambert sum (project_kernel.cu:637-646).
template <int M>
__device__ __forceinline__ void spectral(const SpectrumDev &sp, const float (&ad)[M], float &intensity,
                                         float &photon_prob) {
    float I = 0.0f, P = 0.0f;
    for (int b = 0; b < sp.n_bins; b++) {
        float e = 0.0f;
#pragma unroll
        for (int m = 0; m < M; m++) e += ad[m] * sp.mu[b * M + m];
        const float p = expf(-1.f * e) * sp.pdf[b];
        P += p;
        I += sp.energies[b] * p;
    }
    intensity = I;
    photon_prob = P;
}

// Solid angle of pixel (udx, vdx) (project_kernel.cu:14-133).
__device__ __forceinline__ float solid_angle(const float *W, int udx, int vdx) {
    float cx[4], cy[4], cz[4], cm[4];
    const float cu_off[4] = {0.f, 1.f, 1.f, 0.f};
    const float cv_off[4] = {0.f, 0.f, 1.f, 1.f};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const float cu = udx + cu_off[c];
        const float cv = vdx + cv_off[c];
        cx[c] = cu * W[0] + cv * W[1] + W[2];
        cy[c] = cu * W[3] + cv * W[4] + W[5];
        cz[c] = cu * W[6] + cv * W[7] + W[8];
        cm[c] = sqrtf((cx[c] * cx[c]) + (cy[c] * cy[c]) + (cz[c] * cz[c]));
    }
    const float x02 = (cy[0] * cz[2]) - (cz[0] * cy[2]);
    const float y02 = (cz[0] * cx[2]) - (cx[0] * cz[2]);
    const float z02 = (cx[0] * cy[2]) - (cy[0] * cx[2]);
    const float d01 = (cx[0] * cx[1]) + (cy[0] * cy[1]) + (cz[0] * cz[1]);
    const float d02 = (cx[0] * cx[2]) + (cy[0] * cy[2]) + (cz[0] * cz[2]);
    const float d03 = (cx[0] * cx[3]) + (cy[0] * cy[3]) + (cz[0] * cz[3]);
    const float d12 = (cx[1] * cx[2]) + (cy[1] * cy[2]) + (cz[1] * cz[2]);
    const float d23 = (cx[2] * cx[3]) + (cy[2] * cy[3]) + (cz[2] * cz[3]);
    const float n012 = fabsf((cx[1] * x02) + (cy[1] * y02) + (cz[1] * z02));
    const float n023 = fabsf((cx[3] * x02) + (cy[3] * y02) + (cz[3] * z02));
    const float den012 = (cm[0] * cm[1] * cm[2]) + (d01 * cm[2]) + (d02 * cm[1]) + (d12 * cm[0]);
    const float den023 = (cm[0] * cm[2] * cm[3]) + (d02 * cm[3]) + (d03 * cm[2]) + (d23 * cm[0]);
    float s012 = 2.f * atan2f(n012, den012);
    if (s012 < 0.0f) s012 += 3.141592654f;
    float s023 = 2.f * atan2f(n023, den023);
    if (s023 < 0.0f) s023 += 3.141592654f;
    return s012 + s023;
}

// Tile -> pixel mapping: a 64-lane wave covers an 8x8 pixel tile (a ray
// packet whose samples fall in neighbouring cells at each depth).
__device__ __forceinline__ void lane_pixel(int tile_x, int tile_y, int lane, int &udx, int &vdx) {
    udx = tile_x * 8 + (lane & 7);
    vdx = tile_y * 8 + (lane >> 3);
}

}  // namespace drr
