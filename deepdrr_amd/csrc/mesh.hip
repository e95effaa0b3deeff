// mesh.hip -- triangle meshes in the projector, ray-cast on the GPU.
//
// deepdrr renders meshes with OpenGL (projector.py:1055-1330): dual depth
// peeling of the subtractive meshes of each layer, post-processed per ray by
// peel_postprocess_kernel.cu (tide: cutoff, sort, duplicates, "sea level"),
// and additive blending of distance x facing x density for the additive
// meshes of each (layer, material), with the overlaps of higher subtractive
// layers taken off (density_between.frag); projectKernel then skips the
// volume samples inside a subtractive mesh (project_kernel.cu:363-375,
// 498-521) and adds the additive densities (569-579).
//
// Here a bounding-volume hierarchy over the triangles (built on the host)
// is traversed by one thread per detector pixel, along the same FP32 unit ray
// the march uses, collecting every hit (Moller-Trumbore, FP32, the same
// operation order as the oracle) sorted by (alpha, triangle index).  Per ray
// the kernel then
//   * runs the tide rules on each layer's first max_hits subtractive hits and
//     merges the layers' inside ranges into one sorted list of disjoint
//     ranges (a, b] -- the march tests one range per sample instead of
//     walking every layer's hit list;
//   * sums the additive terms per (layer, material) in FP64 (a sum of
//     +-alpha x density cancels to the chord length x density).
// Entering a mesh (counter-clockwise winding seen from outside) counts +1 in
// a subtractive layer's depth and -1 in the additive facing sum (the z-flip
// of the reference's OpenGL camera, projector.py:866-875, makes entered faces
// back-facing).  A hit repeating the previous kept hit of its list with the
// same facing within 1e-4 mm + 1e-6 alpha is dropped as a duplicate (a ray
// through an edge shared by two triangles; rasterisation counts that point
// once).  At most DRR_MESH_RAY_HITS hits per ray, the nearest, are kept.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "drr_internal.h"
#include "march.h"

using namespace drr;

namespace {

struct BvhNode {
    float lo[3];
    int32_t a;      // leaf: first triangle; inner: index of the right child (the left is the next node)
    float hi[3];
    int32_t count;  // leaf: number of triangles (> 0); inner: 0
};

struct TriDev {
    float v0[3];
    float e1[3];  // v1 - v0 (FP32)
    float e2[3];  // v2 - v0
    int32_t orig; // the caller's triangle index: the tie-break of equal alphas
    int32_t prim;
    int32_t pad;
};

struct PrimDev {
    int32_t material;
    float density;
    int32_t flags;  // bit 0: additive, bit 1: subtractive
    int32_t layer;
};

constexpr int kStack = 64;
constexpr int kSahDepth = 32;  // then median splits: depth <= 32 + 31 < kStack

struct MeshDev {
    const BvhNode *nodes;
    const TriDev *tris;
    const PrimDev *prims;
    int n_layers, max_hits, n_unique;
    uint32_t sub_layers;
    int unique_material[DRR_MAX_MATERIALS];
};

}  // namespace

struct drr_mesh_s {
    drr_ctx *ctx = nullptr;
    int device = 0;
    int n_tris = 0, n_nodes = 0, n_prims = 0, n_layers = 1, max_hits = 32, n_unique = 0;
    uint32_t sub_layers = 0;
    int unique_material[DRR_MAX_MATERIALS] = {};
    BvhNode *d_nodes = nullptr;
    TriDev *d_tris = nullptr;
    PrimDev *d_prims = nullptr;
    // per-view ray terms of drr_project_meshes, grown on demand
    float2 *d_iv = nullptr;
    int32_t *d_cnt = nullptr;
    float *d_add = nullptr;
    int64_t cap_pix = 0;
};

// ---- host: BVH build (binned SAH, 16 bins; median splits below depth 32) ----

namespace {

struct BuildTri {
    float lo[3], hi[3], c[3];
};

struct Builder {
    const std::vector<BuildTri> &bt;
    std::vector<int> &order;
    std::vector<BvhNode> &nodes;

    void bounds(int b, int e, float lo[3], float hi[3], float clo[3], float chi[3]) const {
        for (int a = 0; a < 3; a++) {
            lo[a] = clo[a] = INFINITY;
            hi[a] = chi[a] = -INFINITY;
        }
        for (int i = b; i < e; i++) {
            const BuildTri &t = bt[order[i]];
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], t.lo[a]);
                hi[a] = std::max(hi[a], t.hi[a]);
                clo[a] = std::min(clo[a], t.c[a]);
                chi[a] = std::max(chi[a], t.c[a]);
            }
        }
    }

    static float area(const float lo[3], const float hi[3]) {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return dx * dy + dy * dz + dz * dx;
    }

    int build(int b, int e, int depth) {
        const int idx = (int)nodes.size();
        nodes.push_back(BvhNode{});
        float lo[3], hi[3], clo[3], chi[3];
        bounds(b, e, lo, hi, clo, chi);
        // pad the box so that rounding in the slab test never culls a
        // triangle the ray-triangle test would hit
        for (int a = 0; a < 3; a++) {
            const float pad = 1e-4f * (1.0f + std::max(std::fabs(lo[a]), std::fabs(hi[a])));
            nodes[idx].lo[a] = lo[a] - pad;
            nodes[idx].hi[a] = hi[a] + pad;
        }
        const int n = e - b;
        auto leaf = [&]() {
            nodes[idx].a = b;
            nodes[idx].count = n;
            return idx;
        };
        if (n <= 4) return leaf();
        int axis = 0;
        for (int a = 1; a < 3; a++)
            if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
        const float ext = chi[axis] - clo[axis];
        int mid = -1;
        if (ext > 0 && depth < kSahDepth) {
            constexpr int NB = 16;
            int cnt[NB] = {};
            float blo[NB][3], bhi[NB][3];
            for (int k = 0; k < NB; k++)
                for (int a = 0; a < 3; a++) {
                    blo[k][a] = INFINITY;
                    bhi[k][a] = -INFINITY;
                }
            auto bin_of = [&](int t) {
                int k = (int)((bt[t].c[axis] - clo[axis]) / ext * NB);
                return std::min(NB - 1, std::max(0, k));
            };
            for (int i = b; i < e; i++) {
                const int k = bin_of(order[i]);
                cnt[k]++;
                for (int a = 0; a < 3; a++) {
                    blo[k][a] = std::min(blo[k][a], bt[order[i]].lo[a]);
                    bhi[k][a] = std::max(bhi[k][a], bt[order[i]].hi[a]);
                }
            }
            float best = INFINITY;
            int best_k = -1;
            for (int s = 1; s < NB; s++) {
                float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
                float rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
                int nl = 0, nr = 0;
                for (int k = 0; k < NB; k++) {
                    if (!cnt[k]) continue;
                    float *l = k < s ? llo : rlo, *h = k < s ? lhi : rhi;
                    for (int a = 0; a < 3; a++) {
                        l[a] = std::min(l[a], blo[k][a]);
                        h[a] = std::max(h[a], bhi[k][a]);
                    }
                    (k < s ? nl : nr) += cnt[k];
                }
                if (!nl || !nr) continue;
                const float c = nl * area(llo, lhi) + nr * area(rlo, rhi);
                if (c < best) {
                    best = c;
                    best_k = s;
                }
            }
            if (best_k > 0 && !(best >= n * area(lo, hi) && n <= 16)) {
                mid = (int)(std::partition(order.begin() + b, order.begin() + e,
                                           [&](int t) { return bin_of(t) < best_k; }) - order.begin());
                if (mid == b || mid == e) mid = -1;
            } else if (best_k > 0) {
                return leaf();
            }
        }
        if (mid < 0) {
            if (ext <= 0 && n <= 16) return leaf();
            mid = b + n / 2;
            std::nth_element(order.begin() + b, order.begin() + mid, order.begin() + e,
                             [&](int x, int y) { return bt[x].c[axis] < bt[y].c[axis]; });
        }
        build(b, mid, depth + 1);
        const int right = build(mid, e, depth + 1);
        nodes[idx].a = right;
        nodes[idx].count = 0;
        return idx;
    }
};

void mesh_free(drr_mesh *m) {
    if (!m) return;
    (void)hipFree(m->d_nodes);
    (void)hipFree(m->d_tris);
    (void)hipFree(m->d_prims);
    (void)hipFree(m->d_iv);
    (void)hipFree(m->d_cnt);
    (void)hipFree(m->d_add);
    delete m;
}

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

extern "C" int drr_mesh_create(drr_ctx *ctx, const float *vertices, int n_vertices, const int32_t *triangles,
                               const int32_t *triangle_prim, int n_triangles, const drr_mesh_prim_t *prims,
                               int n_prims, int n_layers, int max_hits, drr_mesh **out) {
    if (!ctx || !out || (n_triangles > 0 && (!vertices || !triangles || !triangle_prim)) || (n_prims > 0 && !prims))
        return drr_set_error(DRR_ERR_INVALID, "drr_mesh_create: NULL argument");
    *out = nullptr;
    if (!ctx->have_scene) return drr_set_error(DRR_ERR_STATE, "drr_mesh_create: scene not set");
    if (n_vertices < 0 || n_triangles < 0 || n_prims < 0 || n_triangles >= (1 << 30))
        return drr_set_error(DRR_ERR_INVALID, "drr_mesh_create: bad counts");
    if (n_layers < 1 || n_layers > DRR_MESH_MAX_LAYERS)
        return drr_set_error(DRR_ERR_INVALID, "drr_mesh_create: n_layers must be in [1, 4]");
    if (max_hits < 4 || max_hits > DRR_MESH_MAX_HITS || max_hits % 4 != 0)
        return drr_set_error(DRR_ERR_INVALID, "drr_mesh_create: max_hits must be a multiple of 4 in [4, 128]");
    const int M = ctx->scene.n_materials;
    for (int p = 0; p < n_prims; p++) {
        if (prims[p].material < 0 || prims[p].material >= M)
            return drr_set_error(DRR_ERR_INVALID, "drr_mesh_create: primitive material out of range");
        if (prims[p].layer < 0 || prims[p].layer >= n_layers)
            return drr_set_error(DRR_ERR_INVALID, "drr_mesh_create: primitive layer out of range");
        if (!std::isfinite(prims[p].density))
            return drr_set_error(DRR_ERR_INVALID, "drr_mesh_create: primitive density must be finite");
    }
    for (int i = 0; i < 3 * n_vertices; i++)
        if (!std::isfinite(vertices[i])) return drr_set_error(DRR_ERR_INVALID, "drr_mesh_create: vertices must be finite");
    std::vector<BuildTri> bt(n_triangles);
    std::vector<TriDev> tris(n_triangles);
    for (int k = 0; k < n_triangles; k++) {
        const int32_t *t = &triangles[3 * k];
        for (int c = 0; c < 3; c++)
            if (t[c] < 0 || t[c] >= n_vertices)
                return drr_set_error(DRR_ERR_INVALID, "drr_mesh_create: vertex index out of range");
        if (triangle_prim[k] < 0 || triangle_prim[k] >= n_prims)
            return drr_set_error(DRR_ERR_INVALID, "drr_mesh_create: triangle primitive out of range");
        const float *v0 = &vertices[3 * t[0]], *v1 = &vertices[3 * t[1]], *v2 = &vertices[3 * t[2]];
        TriDev &d = tris[k];
        for (int a = 0; a < 3; a++) {
            d.v0[a] = v0[a];
            d.e1[a] = v1[a] - v0[a];
            d.e2[a] = v2[a] - v0[a];
            bt[k].lo[a] = std::min(v0[a], std::min(v1[a], v2[a]));
            bt[k].hi[a] = std::max(v0[a], std::max(v1[a], v2[a]));
            bt[k].c[a] = 0.5f * (bt[k].lo[a] + bt[k].hi[a]);
        }
        d.orig = k;
        d.prim = triangle_prim[k];
        d.pad = 0;
    }
    std::vector<int> order(n_triangles);
    for (int k = 0; k < n_triangles; k++) order[k] = k;
    std::vector<BvhNode> nodes;
    if (n_triangles > 0) {
        Builder B{bt, order, nodes};
        B.build(0, n_triangles, 0);
    }
    std::vector<TriDev> tris_sorted(n_triangles);
    for (int i = 0; i < n_triangles; i++) tris_sorted[i] = tris[order[i]];
    std::vector<PrimDev> pd(n_prims);
    drr_mesh *m = new drr_mesh();
    m->ctx = ctx;
    m->device = ctx->device;
    m->n_tris = n_triangles;
    m->n_nodes = (int)nodes.size();
    m->n_prims = n_prims;
    m->n_layers = n_layers;
    m->max_hits = max_hits;
    for (int p = 0; p < n_prims; p++) {
        pd[p] = PrimDev{prims[p].material, prims[p].density,
                        (prims[p].additive ? 1 : 0) | (prims[p].subtractive ? 2 : 0), prims[p].layer};
        if (prims[p].subtractive) m->sub_layers |= 1u << prims[p].layer;
    }
    for (int mat = 0; mat < M; mat++) {
        bool used = false;
        for (int p = 0; p < n_prims; p++) used = used || (prims[p].additive && prims[p].material == mat);
        if (used) m->unique_material[m->n_unique++] = mat;
    }
    DevSwitch ds(ctx->device);
    hipError_t e;
#define MCHK(x)                           \
    do {                                  \
        e = (x);                          \
        if (e != hipSuccess) {            \
            mesh_free(m);                 \
            return drr_hip_error(e, #x);  \
        }                                 \
    } while (0)
    if (m->n_nodes > 0) {
        MCHK(hipMalloc(&m->d_nodes, sizeof(BvhNode) * nodes.size()));
        MCHK(hipMalloc(&m->d_tris, sizeof(TriDev) * tris_sorted.size()));
        MCHK(hipMemcpy(m->d_nodes, nodes.data(), sizeof(BvhNode) * nodes.size(), hipMemcpyHostToDevice));
        MCHK(hipMemcpy(m->d_tris, tris_sorted.data(), sizeof(TriDev) * tris_sorted.size(), hipMemcpyHostToDevice));
    }
    if (n_prims > 0) {
        MCHK(hipMalloc(&m->d_prims, sizeof(PrimDev) * n_prims));
        MCHK(hipMemcpy(m->d_prims, pd.data(), sizeof(PrimDev) * n_prims, hipMemcpyHostToDevice));
    }
#undef MCHK
    *out = m;
    return DRR_OK;
}

extern "C" int drr_mesh_destroy(drr_mesh *mesh) {
    if (!mesh) return DRR_OK;
    DevSwitch ds(mesh->device);
    (void)hipDeviceSynchronize();  // a launch may still read the per-view buffers
    mesh_free(mesh);
    return DRR_OK;
}

// ---- device: per-ray hit collection and mesh terms ------------------------------

namespace {

__device__ __forceinline__ bool ray_triangle(const float o[3], const float d[3], const TriDev &T, float &t_out,
                                             bool &entry) {
    const float pv0 = d[1] * T.e2[2] - d[2] * T.e2[1];
    const float pv1 = d[2] * T.e2[0] - d[0] * T.e2[2];
    const float pv2 = d[0] * T.e2[1] - d[1] * T.e2[0];
    const float det = (T.e1[0] * pv0 + T.e1[1] * pv1) + T.e1[2] * pv2;
    if (det == 0.0f) return false;
    const float inv = 1.0f / det;
    const float tv0 = o[0] - T.v0[0], tv1 = o[1] - T.v0[1], tv2 = o[2] - T.v0[2];
    const float u = ((tv0 * pv0 + tv1 * pv1) + tv2 * pv2) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    const float q0 = tv1 * T.e1[2] - tv2 * T.e1[1];
    const float q1 = tv2 * T.e1[0] - tv0 * T.e1[2];
    const float q2 = tv0 * T.e1[1] - tv1 * T.e1[0];
    const float v = ((d[0] * q0 + d[1] * q1) + d[2] * q2) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    const float t = ((T.e2[0] * q0 + T.e2[1] * q1) + T.e2[2] * q2) * inv;
    if (!(t > 0.0f)) return false;
    t_out = t;
    entry = det > 0.0f;
    return true;
}

__device__ __forceinline__ bool dup_of(float t_kept, float t) { return t - t_kept <= 1e-4f + 1e-6f * t; }

struct RayOut {
    float2 *iv;     // [npix][K]
    int32_t *cnt;   // [npix]
    float *add;     // [n_layers * n_unique][npix]
    int K;
};

__global__ __launch_bounds__(256) void k_mesh_rays(MeshDev m, Geo1 g, float sx, float sy, float sz, float sub_far,
                                                   float add_far, RayOut out, int tiles16_x) {
    const int lb = blockIdx.x;
    const int bx = lb % tiles16_x, by = lb / tiles16_x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int udx, vdx;
    lane_pixel(bx * 2 + (wave & 1), by * 2 + (wave >> 1), lane, udx, vdx);
    if (udx >= g.width || vdx >= g.height) return;
    const int64_t npix = (int64_t)g.width * g.height;
    const int64_t idx = (int64_t)vdx * g.width + udx;
    float rx, ry, rz, L;
    ray_dir(g.W, udx, vdx, rx, ry, rz, L);
    const float o[3] = {sx, sy, sz};
    const float d[3] = {rx, ry, rz};

    // 1. every hit, sorted by (alpha, triangle index); the nearest kept
    float ht[DRR_MESH_RAY_HITS];
    uint32_t hk[DRR_MESH_RAY_HITS];  // triangle index | entry << 31
    int32_t hp[DRR_MESH_RAY_HITS];   // primitive
    int nh = 0;
    if (m.nodes) {
        float inv[3];
        for (int a = 0; a < 3; a++) {
            const float da = fabsf(d[a]) < 1e-20f ? copysignf(1e-20f, d[a]) : d[a];
            inv[a] = 1.0f / da;
        }
        int stack[kStack];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const int ni = stack[--sp];
            const BvhNode nd = m.nodes[ni];
            float tmin = 0.0f, tmax = INFINITY;
            for (int a = 0; a < 3; a++) {
                const float t0 = (nd.lo[a] - o[a]) * inv[a], t1 = (nd.hi[a] - o[a]) * inv[a];
                tmin = fmaxf(tmin, fminf(t0, t1));
                tmax = fminf(tmax, fmaxf(t0, t1));
            }
            if (!(tmin <= tmax)) continue;
            if (nd.count == 0) {
                stack[sp++] = nd.a;
                stack[sp++] = ni + 1;
                continue;
            }
            for (int k = nd.a; k < nd.a + nd.count; k++) {
                const TriDev T = m.tris[k];
                float t;
                bool entry;
                if (!ray_triangle(o, d, T, t, entry)) continue;
                const uint32_t key = (uint32_t)T.orig | (entry ? 0x80000000u : 0u);
                // insert in (t, triangle) order
                int pos = nh;
                while (pos > 0 && (ht[pos - 1] > t || (ht[pos - 1] == t && (hk[pos - 1] & 0x7fffffffu) > (uint32_t)T.orig)))
                    --pos;
                if (pos >= DRR_MESH_RAY_HITS) continue;
                const int last = nh < DRR_MESH_RAY_HITS ? nh : DRR_MESH_RAY_HITS - 1;
                for (int q = last; q > pos; --q) {
                    ht[q] = ht[q - 1];
                    hk[q] = hk[q - 1];
                    hp[q] = hp[q - 1];
                }
                ht[pos] = t;
                hk[pos] = key;
                hp[pos] = T.prim;
                if (nh < DRR_MESH_RAY_HITS) ++nh;
            }
        }
    }

    // 2. per subtractive layer: the tide rules on its first max_hits hits,
    // then its inside ranges (entry, exit] (empty ones kept for the pairing)
    float2 liv[DRR_MESH_MAX_LAYERS][DRR_MESH_MAX_HITS / 2];
    int nl[DRR_MESH_MAX_LAYERS];
    for (int j = 0; j < m.n_layers; j++) {
        nl[j] = 0;
        if (!((m.sub_layers >> j) & 1u)) continue;
        float st[DRR_MESH_MAX_HITS];
        int sf[DRR_MESH_MAX_HITS];
        int n = 0, taken = 0;
        for (int i = 0; i < nh && taken < m.max_hits; i++) {
            const PrimDev p = m.prims[hp[i]];
            if (!(p.flags & 2) || p.layer != j) continue;
            ++taken;  // cut hits still fill the layer's max_hits slots
            const float t = ht[i];
            if (t < 0.00001f || t > sub_far - 0.001f) continue;
            const int f = (hk[i] & 0x80000000u) ? 1 : -1;
            if (n > 0 && sf[n - 1] == f && (st[n - 1] == t || dup_of(st[n - 1], t))) continue;
            st[n] = t;
            sf[n] = f;
            ++n;
        }
        int alt = 0;
        for (int i = 0; i < n; i++) alt += sf[i];
        const int sea = alt > 0 ? alt : 0;
        int prev = 0, cur = 0, depth = 0;
        float start = 0.0f;
        for (int i = 0; i < n; i++) {
            cur = prev + sf[i];
            const bool drop = cur < sea || prev < sea || cur > 1 || prev > 1;
            prev = cur;
            if (drop) continue;
            const int before = depth;
            depth += sf[i];
            if (before <= 0 && depth > 0) {
                start = st[i];
            } else if (before > 0 && depth <= 0) {
                liv[j][nl[j]++] = make_float2(start, st[i]);
            }
        }
        if (depth > 0) liv[j][nl[j]++] = make_float2(start, INFINITY);
    }

    // 3. union over the layers: sorted disjoint ranges (a, b]
    {
        int head[DRR_MESH_MAX_LAYERS] = {0, 0, 0, 0};
        int c = 0;
        float2 curv = make_float2(0.0f, 0.0f);
        bool open = false;
        for (;;) {
            int best = -1;
            for (int j = 0; j < m.n_layers; j++) {
                while (head[j] < nl[j] && !(liv[j][head[j]].x < liv[j][head[j]].y)) ++head[j];  // empty
                if (head[j] < nl[j] && (best < 0 || liv[j][head[j]].x < liv[best][head[best]].x)) best = j;
            }
            if (best < 0) break;
            const float2 r = liv[best][head[best]++];
            if (open && r.x <= curv.y) {
                curv.y = fmaxf(curv.y, r.y);
            } else {
                if (open) out.iv[idx * out.K + c++] = curv;
                curv = r;
                open = true;
            }
        }
        if (open) out.iv[idx * out.K + c++] = curv;
        out.cnt[idx] = c;
    }

    // 4. additive terms per (layer, material): density.frag's
    // distance x facing x density, then density_between.frag's overlaps with
    // each higher subtractive layer's ranges, summed in FP64
    for (int Lr = 0; Lr < m.n_layers; Lr++) {
        for (int u = 0; u < m.n_unique; u++) {
            const int mat = m.unique_material[u];
            double a = 0.0;
            int fsum = 0;
            float last_t = 0.0f;
            int last_mult = 0;
            for (int i = 0; i < nh; i++) {
                const PrimDev p = m.prims[hp[i]];
                if (!(p.flags & 1) || p.layer != Lr || p.material != mat) continue;
                const float t = ht[i];
                if (!(t <= add_far)) continue;
                const int mult = (hk[i] & 0x80000000u) ? -1 : 1;
                if (last_mult == mult && dup_of(last_t, t)) continue;
                last_t = t;
                last_mult = mult;
                const float rho = p.density < 0 ? 0.0f : p.density;
                a += (double)((float)mult * rho) * (double)t;
                fsum += mult;
            }
            for (int S = m.n_layers - 1; S > Lr; S--) {
                if (!((m.sub_layers >> S) & 1u)) continue;
                for (int q = 0; q < nl[S]; q++) {
                    const float nearD = liv[S][q].x, farD = liv[S][q].y;
                    if (!(farD <= add_far) || isinf(farD)) continue;
                    last_mult = 0;
                    for (int i = 0; i < nh; i++) {
                        const PrimDev p = m.prims[hp[i]];
                        if (!(p.flags & 1) || p.layer != Lr || p.material != mat) continue;
                        const float t = ht[i];
                        if (!(t <= add_far)) continue;
                        const int mult = (hk[i] & 0x80000000u) ? -1 : 1;
                        if (last_mult == mult && dup_of(last_t, t)) continue;
                        last_t = t;
                        last_mult = mult;
                        const double rf = (double)((float)mult * p.density);
                        const double dd = t;
                        double r = 0.0;
                        if (dd < nearD) r += -(double)nearD * rf;
                        if (dd < farD) r += (double)farD * rf;
                        if (dd >= nearD && dd < farD) r += -dd * rf;
                        a += r;
                    }
                }
            }
            const float af = (float)a;
            out.add[(int64_t)(Lr * m.n_unique + u) * npix + idx] = fsum == 0 ? fmaxf(af, 0.0f) : 0.0f;
        }
    }
}

// The march's view of one ray's mesh terms (march.h, march_multi).
struct MeshRay {
    const float2 *iv;
    int n, k;
    const float *addv;
    int64_t npix, idx;
    int n_layers, n_unique;
    const int *um;
    __device__ __forceinline__ bool inside(float alpha) {
        while (k < n && alpha > iv[k].y) ++k;
        return k < n && alpha > iv[k].x;
    }
    template <int M>
    __device__ __forceinline__ void add(float (&ad)[M]) {
        for (int u = 0; u < n_unique; u++)
            for (int j = 0; j < n_layers; j++)
                set_m<M>(ad, um[u], get_m<M>(ad, um[u]) + addv[(int64_t)(j * n_unique + u) * npix + idx]);
    }
};

struct MeshTerms {
    const float2 *iv;
    const int32_t *cnt;
    const float *add;
    int K, n_layers, n_unique;
    int unique_material[DRR_MAX_MATERIALS];
};

struct OutMesh {
    float *intensity;
    float *photon_prob;
    float *area_density;
    float *solid_angle;
};

template <int V, int M>
__global__ __launch_bounds__(256) void k_project_mesh(ViewDev vd, SpectrumDev sp, float step, int atten, int air,
                                                      MeshTerms mt, OutMesh out, int tiles16_x) {
    const int lb = blockIdx.x;
    const int bx = lb % tiles16_x, by = lb / tiles16_x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int udx, vdx;
    lane_pixel(bx * 2 + (wave & 1), by * 2 + (wave >> 1), lane, udx, vdx);
    if (udx >= vd.width || vdx >= vd.height) return;
    const int64_t npix = (int64_t)vd.width * vd.height;
    const int64_t idx = (int64_t)vdx * vd.width + udx;
    MeshRay mr{mt.iv + idx * mt.K, mt.cnt[idx], 0, mt.add, npix, idx, mt.n_layers, mt.n_unique, mt.unique_material};
    float ad[M];
    march_multi<V, M, MeshRay>(vd, udx, vdx, step, atten, air, ad, mr);
    float I, P;
    spectral<M>(sp, ad, I, P);
    out.intensity[idx] = I;
    out.photon_prob[idx] = P;
    if (out.area_density) {
#pragma unroll
        for (int m = 0; m < M; m++) out.area_density[m * npix + idx] = ad[m];
    }
    if (out.solid_angle) out.solid_angle[idx] = solid_angle(vd.W, udx, vdx);
}

template <int V>
void launch_mesh_march(int M, dim3 g, dim3 b, hipStream_t st, const ViewDev &vd, const SpectrumDev &sp, float step,
                       int atten, int air, const MeshTerms &mt, const OutMesh &o, int tx) {
    switch (M) {
#define DRR_CASE(MM) \
    case MM: k_project_mesh<V, MM><<<g, b, 0, st>>>(vd, sp, step, atten, air, mt, o, tx); break;
        DRR_CASE(1) DRR_CASE(2) DRR_CASE(3) DRR_CASE(4) DRR_CASE(5) DRR_CASE(6) DRR_CASE(7) DRR_CASE(8) \
        DRR_CASE(9) DRR_CASE(10) DRR_CASE(11) DRR_CASE(12) DRR_CASE(13) DRR_CASE(14) DRR_CASE(15) DRR_CASE(16)
#undef DRR_CASE
    }
}

MeshDev mesh_dev(const drr_mesh *m) {
    MeshDev d{};
    d.nodes = m->d_nodes;
    d.tris = m->d_tris;
    d.prims = m->d_prims;
    d.n_layers = m->n_layers;
    d.max_hits = m->max_hits;
    d.n_unique = m->n_unique;
    d.sub_layers = m->sub_layers;
    for (int i = 0; i < DRR_MAX_MATERIALS; i++) d.unique_material[i] = m->unique_material[i];
    return d;
}

int check_mesh_call(const drr_ctx *ctx, const drr_mesh *mesh, const drr_view_t *view, const float *src,
                    float sdd, const char *who) {
    if (!ctx || !mesh || !view || !src) return drr_set_error(DRR_ERR_INVALID, std::string(who) + ": NULL argument");
    if (mesh->ctx != ctx) return drr_set_error(DRR_ERR_INVALID, std::string(who) + ": mesh belongs to another context");
    if (view->width < 1 || view->height < 1 || view->width > 65536 || view->height > 65536)
        return drr_set_error(DRR_ERR_INVALID, std::string(who) + ": view width/height must be in [1, 65536]");
    if (!std::isfinite(src[0]) || !std::isfinite(src[1]) || !std::isfinite(src[2]) || !std::isfinite(sdd))
        return drr_set_error(DRR_ERR_INVALID, std::string(who) + ": source and distance must be finite");
    return DRR_OK;
}

void launch_rays(const drr_mesh *mesh, const drr_view_t *view, const float *src, float sdd, const RayOut &ro,
                 hipStream_t st) {
    Geo1 g{};
    for (int i = 0; i < 9; i++) g.W[i] = view->world_from_index[i];
    g.width = view->width;
    g.height = view->height;
    const float sub_far = sdd > 0 ? 2.0f * sdd : INFINITY;
    const float add_far = sdd > 0 ? 4.0f * sdd : INFINITY;
    const int tx = (view->width + 15) / 16, ty = (view->height + 15) / 16;
    k_mesh_rays<<<dim3((unsigned)(tx * ty)), dim3(256), 0, st>>>(mesh_dev(mesh), g, src[0], src[1], src[2], sub_far,
                                                                  add_far, ro, tx);
}

}  // namespace

extern "C" int drr_mesh_rays(drr_ctx *ctx, const drr_mesh *mesh, const drr_view_t *view, const float *source_world,
                             float sdd, float *intervals, int32_t *counts, float *additive, int *n_unique,
                             void *stream) {
    int rc = check_mesh_call(ctx, mesh, view, source_world, sdd, "drr_mesh_rays");
    if (rc) return rc;
    if (!intervals || !counts || (!additive && mesh->n_unique > 0))
        return drr_set_error(DRR_ERR_INVALID, "drr_mesh_rays: NULL output");
    if (drr_check_view_finite(view) != DRR_OK) return DRR_ERR_INVALID;
    DevSwitch ds(ctx->device);
    RayOut ro{(float2 *)intervals, counts, additive, mesh->n_layers * mesh->max_hits / 2};
    launch_rays(mesh, view, source_world, sdd, ro, (hipStream_t)stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return drr_hip_error(e, "drr_mesh_rays launch");
    if (n_unique) *n_unique = mesh->n_unique;
    return DRR_OK;
}

extern "C" int drr_project_meshes(drr_ctx *ctx, const drr_volume *const *vols, const drr_view_t *view,
                                  const drr_mesh *mesh_c, const float *source_world, float sdd, float *intensity,
                                  float *photon_prob, float *area_density, float *solid_angle_out,
                                  void *stream) {
    int rc = check_mesh_call(ctx, mesh_c, view, source_world, sdd, "drr_project_meshes");
    if (rc) return rc;
    if (!ctx->have_scene) return drr_set_error(DRR_ERR_STATE, "scene not set (drr_ctx_set_scene)");
    if (ctx->n_bins < 1) return drr_set_error(DRR_ERR_STATE, "spectrum not set (drr_ctx_set_spectrum)");
    if (!intensity || !photon_prob) return drr_set_error(DRR_ERR_INVALID, "drr_project_meshes: NULL output");
    if (view->n_volumes < 0 || view->n_volumes > DRR_MAX_VOLUMES)
        return drr_set_error(DRR_ERR_INVALID, "view.n_volumes must be in [0, 4]");
    for (int i = 0; i < view->n_volumes; i++) {
        if (!vols || !vols[i]) return drr_set_error(DRR_ERR_INVALID, "NULL volume in view");
        if (vols[i]->ctx != ctx) return drr_set_error(DRR_ERR_INVALID, "volume belongs to another context");
    }
    if (drr_check_view_finite(view) != DRR_OK) return DRR_ERR_INVALID;
    drr_mesh *mesh = const_cast<drr_mesh *>(mesh_c);
    DevSwitch ds(ctx->device);
    hipStream_t st = (hipStream_t)stream;
    const int64_t npix = (int64_t)view->width * view->height;
    const int K = mesh->n_layers * mesh->max_hits / 2;
    const int classes = mesh->n_layers * mesh->n_unique;
    if (npix > mesh->cap_pix) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(mesh->d_iv);
        (void)hipFree(mesh->d_cnt);
        (void)hipFree(mesh->d_add);
        mesh->d_iv = nullptr;
        mesh->d_cnt = nullptr;
        mesh->d_add = nullptr;
        mesh->cap_pix = 0;
        hipError_t e1 = hipMalloc(&mesh->d_iv, sizeof(float2) * (size_t)K * npix);
        hipError_t e2 = hipMalloc(&mesh->d_cnt, sizeof(int32_t) * npix);
        hipError_t e3 = classes ? hipMalloc(&mesh->d_add, sizeof(float) * (size_t)classes * npix) : hipSuccess;
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
            (void)hipFree(mesh->d_iv);
            (void)hipFree(mesh->d_cnt);
            (void)hipFree(mesh->d_add);
            mesh->d_iv = nullptr;
            mesh->d_cnt = nullptr;
            mesh->d_add = nullptr;
            return drr_set_error(DRR_ERR_ALLOC, "drr_project_meshes: out of device memory for the ray terms");
        }
        mesh->cap_pix = npix;
    }
    RayOut ro{mesh->d_iv, mesh->d_cnt, mesh->d_add, K};
    (void)hipEventRecord(ctx->ev0, st);
    launch_rays(mesh, view, source_world, sdd, ro, st);
    ViewDev vd;
    view_to_dev(*view, vols, vd);
    SpectrumDev sp{ctx->d_energies, ctx->d_pdf, ctx->d_mu, ctx->n_bins};
    MeshTerms mt{mesh->d_iv, mesh->d_cnt, mesh->d_add, K, mesh->n_layers, mesh->n_unique, {}};
    for (int i = 0; i < DRR_MAX_MATERIALS; i++) mt.unique_material[i] = mesh->unique_material[i];
    OutMesh o{intensity, photon_prob, area_density, solid_angle_out};
    const drr_scene_t &sc = ctx->scene;
    const int tx = (view->width + 15) / 16, ty = (view->height + 15) / 16;
    dim3 grid((unsigned)(tx * ty)), block(256);
    switch (view->n_volumes) {
        case 0: launch_mesh_march<0>(sc.n_materials, grid, block, st, vd, sp, sc.step, sc.attenuate_outside, sc.air_index, mt, o, tx); break;
        case 1: launch_mesh_march<1>(sc.n_materials, grid, block, st, vd, sp, sc.step, sc.attenuate_outside, sc.air_index, mt, o, tx); break;
        case 2: launch_mesh_march<2>(sc.n_materials, grid, block, st, vd, sp, sc.step, sc.attenuate_outside, sc.air_index, mt, o, tx); break;
        case 3: launch_mesh_march<3>(sc.n_materials, grid, block, st, vd, sp, sc.step, sc.attenuate_outside, sc.air_index, mt, o, tx); break;
        default: launch_mesh_march<4>(sc.n_materials, grid, block, st, vd, sp, sc.step, sc.attenuate_outside, sc.air_index, mt, o, tx); break;
    }
    hipError_t e = hipGetLastError();
    (void)hipEventRecord(ctx->ev1, st);
    ctx->timed = true;
    if (e != hipSuccess) return drr_hip_error(e, "drr_project_meshes launch");
    return DRR_OK;
}
