// adjoint.hip -- reverse-mode derivative of the single-volume forward
// projector (drr_project) with respect to the density volume and a rigid
// 6-DoF perturbation of the volume pose.
//
// deepdrr itself is not differentiable (README.md:47); this kernel is the
// exact adjoint of the forward march in march.h (project_kernel.cu:398-646
// semantics): for every ray
//   ad[m]  = step/10 * sum_t wt_t rho(p_t) seg_m(p_t)
//   I      = sum_b E_b pdf_b exp(-sum_m ad[m] mu[b][m]),  P likewise without E_b
// pass 1 re-runs the forward march for ad[] (or reads the area densities the
// forward pass saved), pass 2 re-marches the same samples and scatters
//   dL/dD_corner += wt_t * (sum_d w_d h_d) * w_corner,   h_d = dL/dad_raw[label_d]
// with per-corner sums kept in registers while the ray stays in a cell and
// handed over when it leaves, and accumulates the pose gradient from
//   dL/dp_t = wt_t [ (sum_d w_d h_d) grad(rho) + rho grad(sum_d h_d w_d) ]
// (the second term vanishes in a cell whose labels agree).  Nothing is stored
// per sample; in a cell whose labels agree, pass 2 forms the corner sums over
// the cell's samples at once, in closed form (AdjRay::cell_sums, equal to the
// per-sample sums up to rounding).  drr_adjoint scatters with
// global float atomics; drr_adjoint_jobs runs many views in one persistent
// launch whose workgroups gather the scatter of their 16x16-ray tile in LDS,
// slab by slab along the beam, and add it to the gradient volume once per
// touched voxel.
#include <hip/hip_runtime.h>

#include "drr_internal.h"
#include "jobs.h"
#include "march.h"

using namespace drr;

// d/d(fx, fy, fz) of sum_d v_d w_d for corner values v (corner d = dx + 2dy + 4dz).
__device__ __forceinline__ void grad_trilinear(const float v[8], float fx, float fy, float fz,
                                               float out[3]) {
    const float gx = 1.0f - fx, gy = 1.0f - fy, gz = 1.0f - fz;
    out[0] = (v[1] - v[0]) * (gy * gz) + (v[3] - v[2]) * (fy * gz) + (v[5] - v[4]) * (gy * fz) +
             (v[7] - v[6]) * (fy * fz);
    out[1] = (v[2] - v[0]) * (gx * gz) + (v[3] - v[1]) * (fx * gz) + (v[6] - v[4]) * (gx * fz) +
             (v[7] - v[5]) * (fx * fz);
    out[2] = (v[4] - v[0]) * (gx * gy) + (v[5] - v[1]) * (fx * gy) + (v[6] - v[2]) * (gx * fy) +
             (v[7] - v[3]) * (fx * fy);
}

// ---- where the per-cell corner sums go ----------------------------------------

// Clamped corner voxels of cell (bx, by, bz) (cells in [-1, n-1]^3, voxels < 2^31).
struct CellCorners {
    uint32_t x0, x1, y0, y1, z0, z1;
    __device__ __forceinline__ CellCorners(const VolDev &vol, float bx, float by, float bz) {
        const int cx = clamp_cell(bx, vol.nx), cy = clamp_cell(by, vol.ny), cz = clamp_cell(bz, vol.nz);
        x0 = max(cx, 0);
        x1 = min(cx + 1, vol.nx - 1);
        y0 = max(cy, 0);
        y1 = min(cy + 1, vol.ny - 1);
        z0 = max(cz, 0);
        z1 = min(cz + 1, vol.nz - 1);
    }
};

// Straight to the gradient volume with float atomics.
struct GlobalDep {
    static constexpr bool kDensity = true;
    float *gD;
    // some of the 8 corners (the others are zero in v and skipped per lane)
    __device__ __forceinline__ void part(const VolDev &vol, float bx, float by, float bz, const float (&v)[8]) {
        float a[8];
#pragma unroll
        for (int d = 0; d < 8; d++) a[d] = v[d];
        (*this)(vol, bx, by, bz, a);
    }
    __device__ __forceinline__ void operator()(const VolDev &vol, float bx, float by, float bz, float (&acc)[8]) {
        const CellCorners k(vol, bx, by, bz);
        const uint32_t nx = vol.nx, ny = vol.ny;
        const uint32_t row[4] = {(k.z0 * ny + k.y0) * nx, (k.z0 * ny + k.y1) * nx, (k.z1 * ny + k.y0) * nx,
                                 (k.z1 * ny + k.y1) * nx};
#pragma unroll
        for (int d = 0; d < 8; d++) {
            if (acc[d] != 0.0f) unsafeAtomicAdd(gD + (row[d >> 1] + ((d & 1) ? k.x1 : k.x0)), acc[d]);
            acc[d] = 0.0f;
        }
    }
};

// Nowhere: the pose gradient alone (the density is not differentiated).
struct NullDep {
    static constexpr bool kDensity = false;
    __device__ __forceinline__ void operator()(const VolDev &, float, float, float, float (&acc)[8]) {
#pragma unroll
        for (int d = 0; d < 8; d++) acc[d] = 0.0f;
    }
    __device__ __forceinline__ void part(const VolDev &, float, float, float, const float (&)[8]) {}
};

// Into a workgroup's LDS box [o, o + L) (voxels, per axis); corners outside it
// go to the gradient volume directly, so the box only changes the speed.
struct BoxDep {
    static constexpr bool kDensity = true;
    float *box;
    int ox, oy, oz, Lx, Ly, Lz;
    float *gD;
    // some of the 8 corners: those with a zero sum are skipped lane by lane
    // (an LDS atomic costs per active lane)
    __device__ __forceinline__ void part(const VolDev &vol, float bx, float by, float bz, const float (&v)[8]) {
        float a[8];
#pragma unroll
        for (int d = 0; d < 8; d++) a[d] = v[d];
        deposit<true>(vol, bx, by, bz, a);
    }
    __device__ __forceinline__ void operator()(const VolDev &vol, float bx, float by, float bz, float (&acc)[8]) {
        deposit<false>(vol, bx, by, bz, acc);
    }
    template <bool SKIP_ZERO>
    __device__ __forceinline__ void deposit(const VolDev &vol, float bx, float by, float bz, float (&acc)[8]) {
        const CellCorners k(vol, bx, by, bz);
        const int ix[2] = {(int)k.x0 - ox, (int)k.x1 - ox}, iy[2] = {(int)k.y0 - oy, (int)k.y1 - oy},
                  iz[2] = {(int)k.z0 - oz, (int)k.z1 - oz};
        // x0 <= x1 etc., so the cell is inside when both ends are
        const bool inside = ((uint32_t)ix[0] < (uint32_t)Lx) & ((uint32_t)ix[1] < (uint32_t)Lx) &
                            ((uint32_t)iy[0] < (uint32_t)Ly) & ((uint32_t)iy[1] < (uint32_t)Ly) &
                            ((uint32_t)iz[0] < (uint32_t)Lz) & ((uint32_t)iz[1] < (uint32_t)Lz);
        if (inside) {
            const int row[4] = {(iz[0] * Ly + iy[0]) * Lx, (iz[0] * Ly + iy[1]) * Lx, (iz[1] * Ly + iy[0]) * Lx,
                                (iz[1] * Ly + iy[1]) * Lx};
#pragma unroll
            for (int d = 0; d < 8; d++)
                if (!SKIP_ZERO || acc[d] != 0.0f) atomicAdd(box + (row[d >> 1] + ix[d & 1]), acc[d]);
        } else {
            const uint32_t nx = vol.nx, ny = vol.ny;
#pragma unroll
            for (int d = 0; d < 8; d++) {
                const int a = d & 1, b = (d >> 1) & 1, c = d >> 2;
                if (acc[d] == 0.0f) continue;
                const bool in = ((uint32_t)ix[a] < (uint32_t)Lx) & ((uint32_t)iy[b] < (uint32_t)Ly) &
                                ((uint32_t)iz[c] < (uint32_t)Lz);
                if (in) {
                    atomicAdd(box + ((iz[c] * Ly + iy[b]) * Lx + ix[a]), acc[d]);
                } else {
                    const uint32_t x = a ? k.x1 : k.x0, y = b ? k.y1 : k.y0, z = c ? k.z1 : k.z0;
                    unsafeAtomicAdd(gD + ((z * ny + y) * nx + x), acc[d]);
                }
            }
        }
#pragma unroll
        for (int d = 0; d < 8; d++) acc[d] = 0.0f;
    }
};

// Which cells a call of AdjRay::run may enter, given the floor b of the
// sample position along the slab axis: all, or those up to the end of the
// current slab in march order (sgn * b <= thr; the slab's cell range
// [smin, smax] after clamping to [-1, n-1], folded into one threshold).
struct AllCells {
    __device__ __forceinline__ bool operator()(float) const { return true; }
};
struct UpToSlab {
    float sgn, thr;
    __device__ __forceinline__ UpToSlab(int smin, int smax, bool up, int n_axis) {
        sgn = up ? 1.0f : -1.0f;
        thr = up ? (smax >= n_axis - 1 ? INFINITY : (float)smax) : (smin <= -1 ? INFINITY : -(float)smin);
    }
    __device__ __forceinline__ bool operator()(float b) const { return sgn * b <= thr; }
};

// ---- pass 2 of one ray as a resumable state -------------------------------------

// The second march of one ray: the same samples as march_v1 (same alpha
// sequence, same label cache), scattering G_t w_d(p_t) into the cell's 8
// corner sums, handed to a deposit target whenever the ray leaves a cell
// (and on request), and accumulating the pose sums S0, S1.  `run` stops
// before the first sample whose cell fails a predicate on its coordinate
// along one axis, so a workgroup can march its rays slab by slab.
template <int M, bool POSE>
struct AdjRay {
    RayV1 r;
    Cell c;
    float alpha;
    int t;
    bool live;
    float graw[M];
    float acc[8];   // sum over the cell's samples of G w_d: density gradient of corner d
    float acc1[8];  // sum of (alpha - aref) G w_d (pose only)
    float aref;     // the ray's volume entry: alpha-weighted sums are taken about it
    // (pose kernel) density sums carried into the open cell through the face
    // it was entered by, held apart from acc so the pose term sees the cell's
    // own sums
    float carry[8];
    float hu;       // dL/d(ad_raw) of the material of a cell whose labels agree
    // Pose sums S0 = sum_t dL/dp_t, S1 = sum_t (alpha_t - aref) dL/dp_t, in FP64
    // (thousands of terms of both signs), in registers.  Within a cell, dL/dp_t =
    // G_t grad(rho)(p_t) [+ wt rho_t grad(h)(p_t) if the labels differ, h_d =
    // dL/d(ad_raw) of corner d's material], and grad(rho) along x is
    // sum_yz (D[1yz] - D[0yz]) w_yz, so sum_t G_t grad_x(rho) =
    // sum_yz (D[1yz] - D[0yz]) (acc[0yz] + acc[1yz]) (w_0yz + w_1yz = w_yz):
    // the G term needs no per-sample gradient.  All of this is linear in the
    // sums, so a cell may be flushed in parts.  (They were per-thread LDS
    // slots while the pose kernel ran at 4 waves per SIMD; at 3 the registers
    // are there, and the FP64 adds leave the LDS to the box atomics.)
    double S[6];

    __device__ __forceinline__ void add_S(int k, double v) { S[k] += v; }

    __device__ __forceinline__ void set_h() {
        if (c.uni) hu = get_m<M>(graw, (int)c.l1);
    }

    // h_d of the open cell (a cell whose labels differ), recomputed where needed
    __device__ __forceinline__ void corner_h(float (&h)[8]) const {
#pragma unroll
        for (int d = 0; d < 8; d++) h[d] = get_m<M>(graw, (int)corner_label(c.l0, c.l1, d));
    }

    // Ray setup and the steps before the volume (skip, then the window that
    // re-establishes the label cache), exactly as march_v1.
    __device__ __forceinline__ void begin(const Geo1 &g, const VolDev &vol, int udx, int vdx, float step,
                                          const float (&gr)[M]) {
#pragma unroll
        for (int m = 0; m < M; m++) graw[m] = gr[m];
#pragma unroll
        for (int d = 0; d < 8; d++) {
            acc[d] = 0.0f;
            acc1[d] = 0.0f;
            carry[d] = 0.0f;
        }
        hu = 0.0f;
        if (POSE)
#pragma unroll
            for (int k = 0; k < 6; k++) S[k] = 0.0;
        march_setup_v1(g, udx, vdx, step, r);
        aref = r.mnv;
        live = r.trace;
        alpha = r.minA;
        t = 0;
        if (!live) return;
        const float sx = g.s[0], sy = g.s[1], sz = g.s[2];
        const float rix = r.ri[0], riy = r.ri[1], riz = r.ri[2];
        const float rdom = fmaxf(fabsf(rix), fmaxf(fabsf(riy), fabsf(riz)));
        const float lim = r.mnv - (1.0f / rdom + 2.0f * step);
        skip_steps(r.N, lim, step, alpha, t);
        while (t < r.N && alpha < r.mnv) {
            const float px = sample_pos(sx, alpha, rix);
            const float py = sample_pos(sy, alpha, riy);
            const float pz = sample_pos(sz, alpha, riz);
            const float bx = floorf(px), by = floorf(py), bz = floorf(pz);
            if (!c.have || bx != c.bx || by != c.by || bz != c.bz) c.load(vol, px, py, pz, bx, by, bz);
            alpha += step;
            ++t;
        }
        if (c.have) set_h();
    }

    __device__ __forceinline__ void flush_pose() {
        const float D[8] = {c.d0.x, c.d0.y, c.d0.z, c.d0.w, c.d1.x, c.d1.y, c.d1.z, c.d1.w};
        float g0[3] = {0.f, 0.f, 0.f}, g1[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int b = 1 << k;
#pragma unroll
            for (int d = 0; d < 8; d++) {
                if (d & b) continue;
                const float dD = D[d | b] - D[d];
                g0[k] = fmaf(dD, acc[d] + acc[d | b], g0[k]);
                g1[k] = fmaf(dD, acc1[d] + acc1[d | b], g1[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            add_S(k, (double)g0[k]);
            add_S(k + 3, (double)g1[k]);
        }
#pragma unroll
        for (int d = 0; d < 8; d++) acc1[d] = 0.0f;
    }

    // Hand the open cell's sums to `dep` (the cell stays cached).
    template <class Dep>
    __device__ __forceinline__ void flush(const VolDev &vol, Dep &dep) {
        if (!c.have) return;
        if (POSE) {
            flush_pose();
#pragma unroll
            for (int d = 0; d < 8; d++) {
                acc[d] += carry[d];
                carry[d] = 0.0f;
            }
        }
        dep(vol, c.bx, c.by, c.bz, acc);
    }

    // The ray leaves the open cell through its face normal to axis k (s = +1
    // towards +k, -1 towards -k), into the neighbouring cell.  The 4 corners
    // on that face are corners of the next cell too: their sums stay in
    // registers (moved to the next cell's corner slots) and only the 4 corners
    // left behind are handed to `dep`, lane by lane, so a face crossing costs
    // half the LDS atomic lanes of a full flush.  The density-only kernel
    // carries them in acc itself; the pose kernel, whose pose term needs the
    // cell's own sums, flushes its pose term first and carries them in carry.
    template <class Dep>
    __device__ __forceinline__ void leave_through_face(const VolDev &vol, Dep &dep, int k, int s) {
        const int f = s < 0 ? 1 : 0;  // bit along k of the corners left behind
        // (masks and products instead of selects between array elements, which
        // the compiler would turn into indexed loads from scratch memory)
        const float m0 = k == 0 ? 1.0f : 0.0f, m1 = k == 1 ? 1.0f : 0.0f, m2 = k == 2 ? 1.0f : 0.0f;
        if (POSE) {
            flush_pose();
#pragma unroll
            for (int d = 0; d < 8; d++) acc[d] += carry[d];
        }
        float leave[8], nacc[8];
#pragma unroll
        for (int d = 0; d < 8; d++) {
            const int bit = k == 0 ? (d & 1) : (k == 1 ? ((d >> 1) & 1) : (d >> 2));
            const float on = bit == f ? 1.0f : 0.0f;
            leave[d] = on * acc[d];
            // next cell's corner d is this cell's corner d ^ (1 << k) when on the shared face
            nacc[d] = on * (m0 * acc[d ^ 1] + m1 * acc[d ^ 2] + m2 * acc[d ^ 4]);
        }
        dep.part(vol, c.bx, c.by, c.bz, leave);
#pragma unroll
        for (int d = 0; d < 8; d++) {
            if (POSE) {
                carry[d] = nacc[d];
                acc[d] = 0.0f;
            } else {
                acc[d] = nacc[d];
            }
        }
    }

    // Sums of a one-label cell's n samples of weight 1 in closed form.  The
    // samples sit at u_k = k h (k < n) from the entry sample, whose fractions
    // against the cell are f.  About the run's midpoint u_m = h (n - 1) / 2,
    // with v = u - u_m, corner d = (a, b, c)'s weight is X_a(v) Y_b(v) Z_c(v),
    // X_1 = x_1 + r_x v with x_1 = f_x + r_x u_m and X_0 = 1 - X_1 (so
    // x_0' = -r_x, x_1' = r_x), and likewise along y and z.  The odd central
    // sums of v vanish, so with C2 = sum v^2 = h^2 n (n^2 - 1) / 12 and
    // C4 = sum v^4 = h^4 n (n^2 - 1) (3 n^2 - 7) / 240:
    //   sum_k w_d = n xyz + C2 (x'y'z + x'yz' + xy'z'),
    //   sum_k alpha_k w_d = alpha_m sum_k w_d + C2 (x'yz + xy'z + xyz') + C4 x'y'z',
    // with alpha_m the midpoint's alpha (taken about aref, as acc1 is).  Both are added times hu (the
    // material's dL/d(ad_raw)), as the per-sample loop adds G w_d with G = hu
    // times the weights' sum, which is 1 to rounding.
    __device__ __forceinline__ void cell_sums(int n, float h, float am, float fx, float fy, float fz, float rx,
                                              float ry, float rz) {
        const float nf = (float)n;
        const float um = 0.5f * h * (nf - 1.0f);
        const float x1 = fmaf(rx, um, fx), y1 = fmaf(ry, um, fy), z1 = fmaf(rz, um, fz);
        const float X[2] = {1.0f - x1, x1}, Y[2] = {1.0f - y1, y1}, Z[2] = {1.0f - z1, z1};
        const float Xd[2] = {-rx, rx}, Yd[2] = {-ry, ry}, Zd[2] = {-rz, rz};
        const float h2 = h * h, q = fmaf(nf, nf, -1.0f);
        const float C2 = h2 * nf * q * (1.0f / 12.0f);
        const float C4 = h2 * h2 * nf * q * fmaf(3.0f * nf, nf, -7.0f) * (1.0f / 240.0f);
        const float amid = (am - aref) + um;
        // per (b, c) pair: yz, y'z + yz', y'z'; then each corner (a, b, c)
#pragma unroll
        for (int bc = 0; bc < 4; bc++) {
            const int b = bc & 1, c3 = bc >> 1;
            const float yz = Y[b] * Z[c3], s1 = fmaf(Yd[b], Z[c3], Y[b] * Zd[c3]), qq = Yd[b] * Zd[c3];
#pragma unroll
            for (int a = 0; a < 2; a++) {
                const int d = a + 2 * bc;
                const float t2 = fmaf(Xd[a], s1, X[a] * qq);  // x'(y'z + yz') + x y'z'
                const float W = fmaf(C2, t2, nf * (X[a] * yz));
                acc[d] = fmaf(hu, W, acc[d]);
                if (POSE) {
                    const float t3 = fmaf(X[a], s1, Xd[a] * yz);  // x'yz + x(y'z + yz')
                    const float A = fmaf(amid, W, fmaf(C2, t3, C4 * (Xd[a] * qq)));
                    acc1[d] = fmaf(hu, A, acc1[d]);
                }
            }
        }
    }

    // March on while the floor of the sample position along `axis` of each
    // newly entered cell satisfies go(floor).  The open cell's sums are handed
    // over when the ray leaves it or ends.
    //
    // Cell by cell: every lane still marching hands over its open cell and
    // loads the cell of its current sample together (one execution of that
    // long block for the whole wave), then the wave runs the samples of those
    // cells until every lane's next sample lies outside its own.  Lanes whose
    // cells end early idle for the rest of that inner loop, which costs less
    // than running the hand-over once for each group of lanes that leave
    // together.  Membership is the exact march's: fractions p - b in [0, 1).
    // Loop tests are wave votes and per-lane choices in the inner loop are
    // 0/1 factors (an idle lane's fractions are taken as 0, so its terms are
    // exact zeros): per-lane branches and loop exits cost scalar instructions.
    // Each lane's samples, sums and hand-overs are those of a plain per-sample
    // march.
    template <class Dep, class Go>
    __device__ __forceinline__ void run(const Geo1 &g, const VolDev &vol, float step, int axis, Dep &dep, Go go) {
        const int N = r.N;
        const float sx = g.s[0], sy = g.s[1], sz = g.s[2];
        const float rix = r.ri[0], riy = r.ri[1], riz = r.ri[2];
        // exit distances in steps: fma(-f, j, h) per axis, j = 1 / (r step),
        // h = j going up, 0 going down (j = 0, h = +inf where r = 0)
        const float istep = 1.0f / step;
        const float jx = rix == 0.0f ? 0.0f : __builtin_amdgcn_rcpf(rix) * istep;
        const float jy = riy == 0.0f ? 0.0f : __builtin_amdgcn_rcpf(riy) * istep;
        const float jz = riz == 0.0f ? 0.0f : __builtin_amdgcn_rcpf(riz) * istep;
        const float hx = rix == 0.0f ? INFINITY : (rix < 0.0f ? 0.0f : jx);
        const float hy = riy == 0.0f ? INFINITY : (riy < 0.0f ? 0.0f : jy);
        const float hz = riz == 0.0f ? INFINITY : (riz < 0.0f ? 0.0f : jz);
        bool act = live;
        while (__ballot(act) != 0) {
            // the lane's current sample and its cell
            const float px = sample_pos(sx, alpha, rix), py = sample_pos(sy, alpha, riy),
                        pz = sample_pos(sz, alpha, riz);
            const float bx = floorf(px), by = floorf(py), bz = floorf(pz);
            act = act & (t < N) & (alpha <= r.mxv) & go(axis == 0 ? bx : (axis == 1 ? by : bz));
            if (act && !(bx == c.bx && by == c.by && bz == c.bz)) {
                // one step to a face neighbour (in clamped cell coordinates): carry
                // the shared face; anything else: hand over all 8 corners
                int dd[3] = {0, 0, 0};
                if (Dep::kDensity && c.have) {
                    dd[0] = clamp_cell(bx, vol.nx) - clamp_cell(c.bx, vol.nx);
                    dd[1] = clamp_cell(by, vol.ny) - clamp_cell(c.by, vol.ny);
                    dd[2] = clamp_cell(bz, vol.nz) - clamp_cell(c.bz, vol.nz);
                }
                const bool face = (abs(dd[0]) + abs(dd[1]) + abs(dd[2])) == 1;
                if (face) {
                    const int k = dd[0] ? 0 : (dd[1] ? 1 : 2);
                    leave_through_face(vol, dep, k, dd[0] + dd[1] + dd[2]);
                } else {
                    flush(vol, dep);
                }
                c.load(vol, bx, by, bz);
                set_h();
            }
            // (ex, ey, ez): the fractions of the lane's next sample against the
            // open cell
            float ex = px - c.bx, ey = py - c.by, ez = pz - c.bz;
            // A one-label cell whose samples all have weight 1 and over which
            // the FP32 alpha sequence is linear (its first two increments agree,
            // so no tie alternates, and it stays in one binade) takes its samples
            // at once (cell_sums).  The sample count comes from the exit distance,
            // so a sample within rounding of a face may fall to either cell (the
            // weights on the shared face agree); the last sample is kept for
            // the loop below when the ray ends in the cell.
            bool cf = false;
            if (act && c.uni && t >= 1) {
                const float a1 = alpha + step, s1 = a1 - alpha, s2 = (a1 + step) - a1;
                // the samples before the exit face at the actual increment s1 (a
                // far source moves alpha by up to a quarter step more, or not at all)
                const float nex = fminf(fmaf(-ex, jx, hx), fminf(fmaf(-ey, jy, hy), fmaf(-ez, jz, hz))) *
                                  (s1 > 0.0f ? step / s1 : INFINITY);
                const int n = min(max((int)ceilf(nex), 1), N - 1 - t);
                const float an = fmaf((float)n, s1, alpha);  // alpha after the n samples
                if (n >= 1 && s1 == s2 && (__float_as_uint(an) ^ __float_as_uint(alpha)) < 0x00800000u &&
                    an - s1 <= r.mxv) {
                    cell_sums(n, s1, alpha, ex, ey, ez, rix, riy, riz);
                    alpha = an;
                    t += n;
                    cf = true;
                }
            }
            // the other lanes' samples of the cell, one by one
            bool in = act & !cf;
            if (__ballot(in) == 0) continue;
            do {
                const float gm = in ? 1.0f : 0.0f;
                const float fx = in ? ex : 0.0f, fy = in ? ey : 0.0f, fz = in ? ez : 0.0f;
                float w[8];
                weights8(fx, fy, fz, w);
                // t == 0 or t == N-1 (t < N), as one unsigned comparison
                const float wt = ((uint32_t)(t - 1) >= (uint32_t)(N - 2)) ? 0.5f : 1.0f;
                float sw;
                if (c.uni) {
                    const float ssum = ((((((w[0] + w[1]) + w[2]) + w[3]) + w[4]) + w[5]) + w[6]) + w[7];
                    sw = hu * ssum;
                } else {
                    float h[8];
                    corner_h(h);
                    sw = 0.0f;
#pragma unroll
                    for (int d = 0; d < 8; d++) sw = fmaf(w[d], h[d], sw);
                    if (POSE && in) {
                        // label-gradient term of a cell whose labels differ (rare)
                        const float rho = trilerp(c.d0, c.d1, w);
                        float gh[3];
                        grad_trilinear(h, fx, fy, fz, gh);
                        const float wr = wt * rho;
#pragma unroll
                        for (int k = 0; k < 3; k++) {
                            add_S(k, (double)(wr * gh[k]));
                            add_S(k + 3, (double)((alpha - aref) * wr * gh[k]));
                        }
                    }
                }
                const float G = gm * (wt * sw);
#pragma unroll
                for (int d = 0; d < 8; d++) acc[d] = fmaf(G, w[d], acc[d]);
                if (POSE) {
                    const float aG = (alpha - aref) * G;
#pragma unroll
                    for (int d = 0; d < 8; d++) acc1[d] = fmaf(aG, w[d], acc1[d]);
                }
                alpha = fmaf(gm, step, alpha);
                t += in ? 1 : 0;
                ex = sample_pos(sx, alpha, rix) - c.bx;
                ey = sample_pos(sy, alpha, riy) - c.by;
                ez = sample_pos(sz, alpha, riz) - c.bz;
                in = in & (max(max(__float_as_uint(ex), __float_as_uint(ey)), __float_as_uint(ez)) < 0x3f800000u) &
                     (t < N) & (alpha <= r.mxv);
            } while (__ballot(in) != 0);
        }
        if (live && !(t < N && alpha <= r.mxv)) {
            flush(vol, dep);
            live = false;
        }
    }

    // Pose gradient of the finished ray into gpose.
    __device__ __forceinline__ void end_pose(const Geo1 &g, const float sw_world[3], const float center[3],
                                             double (&gpose)[6]) const {
        if (!r.trace) return;
        // p = A x + b - 1 with x = s_w + alpha r_w; perturbed x' = R(w)^T (x - c - t) + c:
        // dp/dt = -A, dp/dw = A [x - c]_x  ->  dL/dt = -A^T S0,
        // dL/dw = (A^T S0) x (s_w + aref r_w - c) + (A^T S1) x r_w with S1 summed
        // over alpha - aref (a far source's alpha is thousands of times the
        // volume's size; the two terms would cancel to a few digits in FP32)
        const float *A = g.A;
        const double S0[3] = {S[0], S[1], S[2]}, S1[3] = {S[3], S[4], S[5]};
        double V0[3], V1[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            V0[j] = (double)A[j] * S0[0] + (double)A[4 + j] * S0[1] + (double)A[8 + j] * S0[2];
            V1[j] = (double)A[j] * S1[0] + (double)A[4 + j] * S1[1] + (double)A[8 + j] * S1[2];
        }
        const double rw[3] = {r.rx, r.ry, r.rz};
        const double q[3] = {(double)sw_world[0] - center[0] + (double)aref * rw[0],
                             (double)sw_world[1] - center[1] + (double)aref * rw[1],
                             (double)sw_world[2] - center[2] + (double)aref * rw[2]};
        gpose[0] -= V0[0];
        gpose[1] -= V0[1];
        gpose[2] -= V0[2];
        gpose[3] += (V0[1] * q[2] - V0[2] * q[1]) + (V1[1] * rw[2] - V1[2] * rw[1]);
        gpose[4] += (V0[2] * q[0] - V0[0] * q[2]) + (V1[2] * rw[0] - V1[0] * rw[2]);
        gpose[5] += (V0[0] * q[1] - V0[1] * q[0]) + (V1[0] * rw[1] - V1[1] * rw[0]);
    }
};

// dL/d(ad_raw[m]) of one pixel from its area densities ad[m] (g/cm^2) for
// output gradients gi (intensity) and gp (photon_prob); ad_raw = ad * 10 / step.
template <int M>
__device__ __forceinline__ void graw_from_ad(const SpectrumDev &sp, const float (&ad)[M], float step, float gi,
                                             float gp, float (&graw)[M]) {
    float gad[M];
#pragma unroll
    for (int m = 0; m < M; m++) gad[m] = 0.0f;
    for (int b = 0; b < sp.n_bins; b++) {
        float e = 0.0f;
#pragma unroll
        for (int m = 0; m < M; m++) e += ad[m] * sp.mu[b * M + m];
        const float coef = (gi * sp.energies[b] + gp) * sp.pdf[b] * expf(-e);
#pragma unroll
        for (int m = 0; m < M; m++) gad[m] -= coef * sp.mu[b * M + m];
    }
#pragma unroll
    for (int m = 0; m < M; m++) graw[m] = gad[m] * step / 10.0f;
}

// Pass 1 of one pixel: the forward march, then graw.
template <int M, bool ATTEN>
__device__ __forceinline__ void backward_spectral(const Geo1 &g, const VolDev &vol, const SpectrumDev &sp,
                                                  float step, int air, int udx, int vdx, float gi, float gp,
                                                  float (&graw)[M]) {
    float ad[M];
    march_v1<M, ATTEN>(g, vol, udx, vdx, step, air, ad);
    graw_from_ad<M>(sp, ad, step, gi, gp, graw);
}

// Both passes for one pixel, scattering straight into the gradient volume.
template <int M, bool ATTEN, bool POSE>
__device__ __forceinline__ void adjoint_ray(const Geo1 &g, const VolDev &vol, const SpectrumDev &sp, float step,
                                            int air, int udx, int vdx, float gi, float gp, const float sw[3],
                                            const float center[3], float *gD, double (&gpose)[6]) {
    if (gi == 0.0f && gp == 0.0f) return;
    float graw[M];
    backward_spectral<M, ATTEN>(g, vol, sp, step, air, udx, vdx, gi, gp, graw);
    AdjRay<M, POSE> ray;
    ray.begin(g, vol, udx, vdx, step, graw);
    GlobalDep dep{gD};
    ray.run(g, vol, step, 0, dep, AllCells{});
    if (POSE) ray.end_pose(g, sw, center, gpose);
}

__device__ __forceinline__ void flush_pose_wave(double (&gpose)[6], float *gPose) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        double v = gpose[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0 && v != 0.0) unsafeAtomicAdd(gPose + k, (float)v);
        gpose[k] = 0.0;
    }
}

struct AdjArgs {
    const float *gI;  // [H][W] (or job images at out_offset), may be NULL
    const float *gP;  // may be NULL
    const float *AD;  // saved forward area densities (M planes per job at M * out_offset), may be NULL
    float c[3];       // pose centre in world
    float *gD;        // [nz][ny][nx]
    float *gPose;     // [6] or NULL
};

template <int M, bool ATTEN, bool POSE>
__global__ __launch_bounds__(256) void k_adjoint_v1(Geo1 g, VolDev vol, SpectrumDev sp, float step, int air,
                                                    AdjArgs a, float swx, float swy, float swz, int tiles16_x) {
    // tiles in row-major order, dispatched round-robin over the XCDs, so
    // every XCD gets a share of every band of rows (of ray lengths)
    const int lb = blockIdx.x;
    const int bx = lb % tiles16_x, by = lb / tiles16_x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int udx, vdx;
    lane_pixel(bx * 2 + (wave & 1), by * 2 + (wave >> 1), lane, udx, vdx);
    double gpose[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (udx < g.width && vdx < g.height) {
        const int64_t idx = (int64_t)vdx * g.width + udx;
        const float sw[3] = {swx, swy, swz};
        adjoint_ray<M, ATTEN, POSE>(g, vol, sp, step, air, udx, vdx, a.gI ? a.gI[idx] : 0.0f,
                                    a.gP ? a.gP[idx] : 0.0f, sw, a.c, a.gD, gpose);
    }
    if (POSE) flush_pose_wave(gpose, a.gPose);
}

__global__ void k_pose_zero(double *acc) {
    if (threadIdx.x < 6) acc[threadIdx.x] = 0.0;
}

__global__ void k_pose_finish(const double *acc, float *gPose) {
    if (threadIdx.x < 6) gPose[threadIdx.x] += (float)acc[threadIdx.x];
}

// ---- batched adjoint: workgroup tiles, LDS accumulation slab by slab --------------

// A workgroup takes a 16x16-pixel tile of one view (4 waves of 8x8 rays).
// Its rays form a narrow beam; the beam is cut into slabs of ADJ_SLAB cells
// along the axis its central ray advances fastest on, and all rays march one
// slab at a time.  Corner sums go into an LDS box covering the slab
// (ADJ_SLAB + 1 voxel planes along that axis, ADJ_BOX voxels across, centred
// on the central ray) with LDS atomics; after each slab the box is added to
// the gradient volume once per touched voxel.  Corners outside the box (a
// beam wider than the box, a ray against the slab order) go to the gradient
// volume directly, so results do not depend on the geometry, only the
// number of global atomics does.
constexpr int ADJ_TILE = 16;
constexpr int ADJ_SLAB = 15;
constexpr int ADJ_BOX = 20;
constexpr int ADJ_BOX_CELLS = (ADJ_SLAB + 1) * ADJ_BOX * ADJ_BOX;

// World ray through the detector point (fu, fv) (pixel-edge coordinates).
__device__ __forceinline__ void ray_through(const float *W, float fu, float fv, float r[3]) {
    r[0] = fu * W[0] + fv * W[1] + W[2];
    r[1] = fu * W[3] + fv * W[4] + W[5];
    r[2] = fu * W[6] + fv * W[7] + W[8];
}

// Occupancy targets that fit without (or, for more materials, with a few)
// register spills: 5 waves per SIMD for the density gradient alone, 3 with
// the pose sums (density, pose moments and the carried face: 4 waves spill in
// the cell loop and ran 4 % slower).
template <int POSE>
struct AdjOcc {
    static constexpr int waves = POSE ? 3 : 5;
};

template <int M, bool ATTEN, bool POSE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(AdjOcc<POSE>::waves))) void k_adjoint_jobs(const JobDev *__restrict__ jobs, const int2 *__restrict__ items,
                                                      const int *__restrict__ seg, int *counters,
                                                      const float *__restrict__ sw_jobs, VolDev vol, SpectrumDev sp,
                                                      float step, int air, AdjArgs a, double *pose64) {
    __shared__ float box[ADJ_BOX_CELLS];
    __shared__ double s_pose[4][6];
    for (int i = threadIdx.x; i < ADJ_BOX_CELLS; i += blockDim.x) box[i] = 0.0f;
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for_each_block_item(jobs, items, seg, counters, [&](int j, const JobDev &J, int tile) __attribute__((always_inline)) {
        const int tx = tile % J.tiles_x, ty = tile / J.tiles_x;
        int udx, vdx;
        lane_pixel(tx * 2 + (wave & 1), ty * 2 + (wave >> 1), lane, udx, vdx);
        float gi = 0.0f, gp = 0.0f;
        if (udx < J.g.width && vdx < J.g.height) {
            const int64_t o = J.out_offset + (int64_t)vdx * J.g.width + udx;
            gi = a.gI ? a.gI[o] : 0.0f;
            gp = a.gP ? a.gP[o] : 0.0f;
        }
        AdjRay<M, POSE> ray;
        ray.live = false;
        ray.r.trace = false;
        if (gi != 0.0f || gp != 0.0f) {
            float graw[M];
            if (a.AD) {
                // the forward pass's area densities: no need to march twice
                const int64_t npix = (int64_t)J.g.width * J.g.height;
                const float *ad_p = a.AD + M * J.out_offset + (int64_t)vdx * J.g.width + udx;
                float ad[M];
#pragma unroll
                for (int m = 0; m < M; m++) ad[m] = ad_p[m * npix];
                graw_from_ad<M>(sp, ad, step, gi, gp, graw);
            } else {
                backward_spectral<M, ATTEN>(J.g, vol, sp, step, air, udx, vdx, gi, gp, graw);
            }
            ray.begin(J.g, vol, udx, vdx, step, graw);
        }
        // The tile's central ray (the same value in every thread) fixes the
        // slab axis, the slab order and the box centre.
        float rc[3], rci[3];
        ray_through(J.g.W, (float)(tx * ADJ_TILE) + 0.5f * ADJ_TILE, (float)(ty * ADJ_TILE) + 0.5f * ADJ_TILE, rc);
#pragma unroll
        for (int k = 0; k < 3; k++) rci[k] = J.g.A[4 * k] * rc[0] + J.g.A[4 * k + 1] * rc[1] + J.g.A[4 * k + 2] * rc[2];
        // (selects, not array indexing by `axis`, so nothing lands in scratch memory)
        int axis = fabsf(rci[1]) > fabsf(rci[0]) ? 1 : 0;
        float ra = axis ? rci[1] : rci[0];
        if (fabsf(rci[2]) > fabsf(ra)) {
            axis = 2;
            ra = rci[2];
        }
        const bool up = !(ra < 0.0f);
        const int n_axis = axis == 0 ? vol.nx : (axis == 1 ? vol.ny : vol.nz);
        const float s_axis = axis == 0 ? J.g.s[0] : (axis == 1 ? J.g.s[1] : J.g.s[2]);
        const int n_slabs = (n_axis + 1 + ADJ_SLAB - 1) / ADJ_SLAB;  // cells -1 .. n-1
        for (int k = 0; k < n_slabs; k++) {
            // cells [smin, smax] along the axis; their corners lie in [smin, smax + 1]
            const int smin = up ? -1 + k * ADJ_SLAB : n_axis - 1 - (k + 1) * ADJ_SLAB + 1;
            const int smax = smin + ADJ_SLAB - 1;
            // box across the beam: centred on the central ray at the slab's middle plane
            float o[3] = {0.f, 0.f, 0.f};
            if (ra != 0.0f) {
                const float mid = 0.5f * (float)(smin + smax + 1);
                const float al = (mid + 1.0f - s_axis) / ra;
#pragma unroll
                for (int q = 0; q < 3; q++) o[q] = J.g.s[q] + al * rci[q] - 1.0f;
            }
            int org[3], len[3];
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const float f = __builtin_amdgcn_fmed3f(floorf(o[q]), -1048576.0f, 1048576.0f);
                org[q] = q == axis ? smin : (int)f - ADJ_BOX / 2 + 1;
                len[q] = q == axis ? ADJ_SLAB + 1 : ADJ_BOX;
            }
            BoxDep dep{box, org[0], org[1], org[2], len[0], len[1], len[2], a.gD};
            ray.run(J.g, vol, step, axis, dep, UpToSlab(smin, smax, up, n_axis));
            if (ray.live) ray.flush(vol, dep);
            __syncthreads();
            // add the box to the gradient volume and clear it
            const float inv_x = 1.0f / (float)len[0], inv_y = 1.0f / (float)len[1];
            const uint32_t nx = vol.nx, ny = vol.ny;
            for (int i = threadIdx.x; i < ADJ_BOX_CELLS; i += blockDim.x) {
                const float v = box[i];
                if (v != 0.0f) {
                    const int iyz = (int)(((float)i + 0.5f) * inv_x);
                    const int ix = i - iyz * len[0];
                    const int iz = (int)(((float)iyz + 0.5f) * inv_y);
                    const int iy = iyz - iz * len[1];
                    const int x = org[0] + ix, y = org[1] + iy, z = org[2] + iz;
                    if (x >= 0 && x < vol.nx && y >= 0 && y < vol.ny && z >= 0 && z < vol.nz)
                        unsafeAtomicAdd(a.gD + (((uint32_t)z * ny + (uint32_t)y) * nx + (uint32_t)x), v);
                    box[i] = 0.0f;
                }
            }
            __syncthreads();
        }
        // anything left (none when every ray follows the slab order) goes straight out
        GlobalDep gdep{a.gD};
        ray.run(J.g, vol, step, axis, gdep, AllCells{});
        if (POSE) {
            double gpose[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            const float sw[3] = {sw_jobs[3 * j], sw_jobs[3 * j + 1], sw_jobs[3 * j + 2]};
            ray.end_pose(J.g, sw, a.c, gpose);
#pragma unroll
            for (int q = 0; q < 6; q++) {
                double v = gpose[q];
                for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
                if (lane == 0) s_pose[wave][q] = v;
            }
            __syncthreads();
            if (threadIdx.x < 6) {
                const double v = s_pose[0][threadIdx.x] + s_pose[1][threadIdx.x] + s_pose[2][threadIdx.x] +
                                 s_pose[3][threadIdx.x];
                if (v != 0.0) unsafeAtomicAdd(pose64 + threadIdx.x, v);
            }
            __syncthreads();
        }
    });
}

// Pose gradient alone (registration with a fixed density): the same rays and
// pose sums as k_adjoint_jobs<..., true>, without any density scatter, so no
// slabs or LDS box.
template <int M, bool ATTEN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_adjoint_pose_jobs(const JobDev *__restrict__ jobs,
                                                           const int2 *__restrict__ items,
                                                           const int *__restrict__ seg, int *counters,
                                                           const float *__restrict__ sw_jobs, VolDev vol,
                                                           SpectrumDev sp, float step, int air, AdjArgs a,
                                                           double *pose64) {
    __shared__ double s_pose[4][6];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for_each_block_item(jobs, items, seg, counters, [&](int j, const JobDev &J, int tile) __attribute__((always_inline)) {
        const int tx = tile % J.tiles_x, ty = tile / J.tiles_x;
        int udx, vdx;
        lane_pixel(tx * 2 + (wave & 1), ty * 2 + (wave >> 1), lane, udx, vdx);
        double gpose[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        float gi = 0.0f, gp = 0.0f;
        if (udx < J.g.width && vdx < J.g.height) {
            const int64_t o = J.out_offset + (int64_t)vdx * J.g.width + udx;
            gi = a.gI ? a.gI[o] : 0.0f;
            gp = a.gP ? a.gP[o] : 0.0f;
        }
        if (gi != 0.0f || gp != 0.0f) {
            float graw[M];
            if (a.AD) {
                const int64_t npix = (int64_t)J.g.width * J.g.height;
                const float *ad_p = a.AD + M * J.out_offset + (int64_t)vdx * J.g.width + udx;
                float ad[M];
#pragma unroll
                for (int m = 0; m < M; m++) ad[m] = ad_p[m * npix];
                graw_from_ad<M>(sp, ad, step, gi, gp, graw);
            } else {
                backward_spectral<M, ATTEN>(J.g, vol, sp, step, air, udx, vdx, gi, gp, graw);
            }
            AdjRay<M, true> ray;
            ray.begin(J.g, vol, udx, vdx, step, graw);
            NullDep dep;
            ray.run(J.g, vol, step, 0, dep, AllCells{});
            const float sw[3] = {sw_jobs[3 * j], sw_jobs[3 * j + 1], sw_jobs[3 * j + 2]};
            ray.end_pose(J.g, sw, a.c, gpose);
        }
#pragma unroll
        for (int q = 0; q < 6; q++) {
            double v = gpose[q];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (lane == 0) s_pose[wave][q] = v;
        }
        __syncthreads();
        if (threadIdx.x < 6) {
            const double v = s_pose[0][threadIdx.x] + s_pose[1][threadIdx.x] + s_pose[2][threadIdx.x] +
                             s_pose[3][threadIdx.x];
            if (v != 0.0) unsafeAtomicAdd(pose64 + threadIdx.x, v);
        }
        __syncthreads();
    });
}

template <int M, bool ATTEN>
static void launch_adj(bool pose, dim3 gr, dim3 bl, hipStream_t st, const Geo1 &g, const VolDev &v,
                       const SpectrumDev &sp, float step, int air, const AdjArgs &a, const float *sw, int tx) {
    if (pose)
        k_adjoint_v1<M, ATTEN, true><<<gr, bl, 0, st>>>(g, v, sp, step, air, a, sw[0], sw[1], sw[2], tx);
    else
        k_adjoint_v1<M, ATTEN, false><<<gr, bl, 0, st>>>(g, v, sp, step, air, a, sw[0], sw[1], sw[2], tx);
}

template <int M, bool ATTEN>
static void launch_adj_jobs(bool pose, const JobTable &t, hipStream_t st, const float *sw_jobs, const VolDev &v,
                            const SpectrumDev &sp, float step, int air, const AdjArgs &a, double *pose64) {
    dim3 gr(t.grid), bl(t.block);
    if (!a.gD)
        k_adjoint_pose_jobs<M, ATTEN><<<gr, bl, 0, st>>>(t.d_jobs, t.d_items, t.d_seg, t.d_counters, sw_jobs, v, sp,
                                                         step, air, a, pose64);
    else if (pose)
        k_adjoint_jobs<M, ATTEN, true><<<gr, bl, 0, st>>>(t.d_jobs, t.d_items, t.d_seg, t.d_counters, sw_jobs, v,
                                                          sp, step, air, a, pose64);
    else
        k_adjoint_jobs<M, ATTEN, false><<<gr, bl, 0, st>>>(t.d_jobs, t.d_items, t.d_seg, t.d_counters, sw_jobs, v,
                                                           sp, step, air, a, pose64);
}

static int check_adj(drr_ctx *ctx, const drr_volume *vol, const float *gI, const float *gP, const float *gD,
                     const float *gPose, const float *sw, const float *c, bool pose_only_ok = false) {
    if (!ctx || !vol) return drr_set_error(DRR_ERR_INVALID, "drr_adjoint: NULL context or volume");
    if (vol->ctx != ctx) return drr_set_error(DRR_ERR_INVALID, "drr_adjoint: volume of another context");
    if (!ctx->have_scene || ctx->n_bins < 1) return drr_set_error(DRR_ERR_STATE, "drr_adjoint: scene/spectrum not set");
    if (!gD && !(pose_only_ok && gPose))
        return drr_set_error(DRR_ERR_INVALID, pose_only_ok ? "drr_adjoint_jobs: neither grad_density nor grad_pose given"
                                                           : "drr_adjoint: grad_density is NULL");
    if (!gI && !gP) return drr_set_error(DRR_ERR_INVALID, "drr_adjoint: no output gradient given");
    if (gPose && (!sw || !c))
        return drr_set_error(DRR_ERR_INVALID, "drr_adjoint: pose gradient needs source_world and pose_center_world");
    return DRR_OK;
}

extern "C" int drr_adjoint(drr_ctx *ctx, const drr_volume *vol, const drr_view_t *view,
                           const float *grad_intensity, const float *grad_photon_prob,
                           const float *source_world, const float *pose_center_world,
                           float *grad_density, float *grad_pose, void *stream) {
    // before drr_check_view, which reads n_volumes entries of the volume array
    if (view && view->n_volumes != 1) return drr_set_error(DRR_ERR_UNSUPPORTED, "drr_adjoint: one volume per view");
    const drr_volume *vols[1] = {vol};
    int rc = drr_check_view(ctx, vols, view);
    if (rc) return rc;
    rc = check_adj(ctx, vol, grad_intensity, grad_photon_prob, grad_density, grad_pose, source_world,
                   pose_center_world);
    if (rc) return rc;
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != ctx->device) DRR_HIP_CHECK(hipSetDevice(ctx->device));
    Geo1 g;
    view_to_geo1(*view, g);
    const VolDev v = vol_to_dev(vol);
    SpectrumDev sp{ctx->d_energies, ctx->d_pdf, ctx->d_mu, ctx->n_bins};
    AdjArgs a{};
    a.gI = grad_intensity;
    a.gP = grad_photon_prob;
    float sw[3] = {0.f, 0.f, 0.f};
    for (int i = 0; i < 3; i++) {
        sw[i] = source_world ? source_world[i] : 0.0f;
        a.c[i] = pose_center_world ? pose_center_world[i] : 0.0f;
    }
    a.gD = grad_density;
    a.gPose = grad_pose;
    const drr_scene_t &sc = ctx->scene;
    const int tx = (view->width + 15) / 16, ty = (view->height + 15) / 16;
    dim3 grid((unsigned)(tx * ty)), block(256);
    hipStream_t st = (hipStream_t)stream;
    const bool pose = grad_pose != nullptr;
    (void)hipEventRecord(ctx->ev0, st);
    switch (sc.n_materials) {
#define DRR_CASE(MM)                                                                              \
    case MM:                                                                                      \
        if (sc.attenuate_outside)                                                                 \
            launch_adj<MM, true>(pose, grid, block, st, g, v, sp, sc.step, sc.air_index, a, sw, tx);  \
        else                                                                                      \
            launch_adj<MM, false>(pose, grid, block, st, g, v, sp, sc.step, sc.air_index, a, sw, tx); \
        break;
        DRR_CASE(1) DRR_CASE(2) DRR_CASE(3) DRR_CASE(4) DRR_CASE(5) DRR_CASE(6) DRR_CASE(7) DRR_CASE(8) \
        DRR_CASE(9) DRR_CASE(10) DRR_CASE(11) DRR_CASE(12) DRR_CASE(13) DRR_CASE(14) DRR_CASE(15) DRR_CASE(16)
#undef DRR_CASE
    }
    hipError_t e = hipGetLastError();
    (void)hipEventRecord(ctx->ev1, st);
    ctx->timed = true;
    if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
    if (e != hipSuccess) return drr_hip_error(e, "drr_adjoint launch");
    return DRR_OK;
}

extern "C" int drr_adjoint_jobs(drr_ctx *ctx, const drr_volume *vol, const drr_job_t *jobs, int n_jobs,
                                const float *grad_intensity, const float *grad_photon_prob,
                                const float *source_world, const float *pose_center_world,
                                const float *area_density, float *grad_density, float *grad_pose, void *stream) {
    int rc = check_adj(ctx, vol, grad_intensity, grad_photon_prob, grad_density, grad_pose, source_world,
                       pose_center_world, true);
    if (rc) return rc;
    if (!jobs || n_jobs < 1) return drr_set_error(DRR_ERR_INVALID, "drr_adjoint_jobs: no jobs");
    for (int j = 0; j < n_jobs; j++)
        if (jobs[j].volume_index != 0)
            return drr_set_error(DRR_ERR_INVALID, "drr_adjoint_jobs: every job must use volume 0");
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != ctx->device) DRR_HIP_CHECK(hipSetDevice(ctx->device));
    const drr_volume *vols[1] = {vol};
    JobTable t;
    // 16x16-pixel workgroup tiles, 4 waves per workgroup; as many persistent
    // workgroups per CU as the kernel's occupancy target holds (LDS allows it)
    const bool pose = grad_pose != nullptr;
    const int blocks_per_cu = pose ? AdjOcc<1>::waves : AdjOcc<0>::waves;  // pose only: 4, as with density
    const drr_batch_tuning_t tune{ADJ_TILE, ADJ_TILE, 4, blocks_per_cu};
    rc = drr_jobs_build_tiles(ctx, vols, 1, jobs, n_jobs, tune, t);
    if (rc != DRR_OK) return rc;
    // scratch: FP64 pose accumulator [6], then the job sources [3 * n_jobs]
    double *d_pose64 = nullptr;
    hipError_t e = hipMalloc(&d_pose64, sizeof(double) * 6 + sizeof(float) * 3 * n_jobs);
    float *d_sw = e == hipSuccess ? reinterpret_cast<float *>(d_pose64 + 6) : nullptr;
    if (e == hipSuccess && source_world)
        e = hipMemcpy(d_sw, source_world, sizeof(float) * 3 * n_jobs, hipMemcpyHostToDevice);
    else if (e == hipSuccess)
        e = hipMemset(d_sw, 0, sizeof(float) * 3 * n_jobs);
    if (e != hipSuccess) {
        drr_jobs_free(t);
        (void)hipFree(d_pose64);
        return drr_hip_error(e, "drr_adjoint_jobs: scratch");
    }
    const VolDev v = vol_to_dev(vol);
    SpectrumDev sp{ctx->d_energies, ctx->d_pdf, ctx->d_mu, ctx->n_bins};
    AdjArgs a{};
    a.gI = grad_intensity;
    a.gP = grad_photon_prob;
    a.AD = area_density;
    for (int i = 0; i < 3; i++) a.c[i] = pose_center_world ? pose_center_world[i] : 0.0f;
    a.gD = grad_density;
    a.gPose = grad_pose;
    const drr_scene_t &sc = ctx->scene;
    hipStream_t st = (hipStream_t)stream;
    (void)hipEventRecord(ctx->ev0, st);
    k_jobs_reset<<<1, 64, 0, st>>>(t.d_counters);
    if (pose) k_pose_zero<<<1, 64, 0, st>>>(d_pose64);
    switch (sc.n_materials) {
#define DRR_CASE(MM)                                                                                  \
    case MM:                                                                                          \
        if (sc.attenuate_outside)                                                                     \
            launch_adj_jobs<MM, true>(pose, t, st, d_sw, v, sp, sc.step, sc.air_index, a, d_pose64);  \
        else                                                                                          \
            launch_adj_jobs<MM, false>(pose, t, st, d_sw, v, sp, sc.step, sc.air_index, a, d_pose64); \
        break;
        DRR_CASE(1) DRR_CASE(2) DRR_CASE(3) DRR_CASE(4) DRR_CASE(5) DRR_CASE(6) DRR_CASE(7) DRR_CASE(8) \
        DRR_CASE(9) DRR_CASE(10) DRR_CASE(11) DRR_CASE(12) DRR_CASE(13) DRR_CASE(14) DRR_CASE(15) DRR_CASE(16)
#undef DRR_CASE
    }
    if (pose) k_pose_finish<<<1, 64, 0, st>>>(d_pose64, grad_pose);
    e = hipGetLastError();
    (void)hipEventRecord(ctx->ev1, st);
    ctx->timed = true;
    hipError_t e2 = hipStreamSynchronize(st);
    drr_jobs_free(t);
    (void)hipFree(d_pose64);
    if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
    if (e != hipSuccess) return drr_hip_error(e, "drr_adjoint_jobs launch");
    if (e2 != hipSuccess) return drr_hip_error(e2, "drr_adjoint_jobs");
    return DRR_OK;
}
