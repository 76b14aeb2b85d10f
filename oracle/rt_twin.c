/*
 * rt_twin.c — ORACLE / TEST INFRASTRUCTURE ONLY.
 *
 * A CPU (C11 + OpenMP) restatement of the ray-casting path that librtx_hip.so implements on the
 * GPU, plus an independent FP64 closest-hit oracle. Only tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py load this library; the product never does.
 *
 * The reference raycaster is a stub (rendering/_raycaster.py:30-36: `_build_ads` counts triangles,
 * `ray_cast` is `pass`), so there is no reference algorithm to transcribe. What this file follows
 * from the reference is its API and conventions:
 *   - triangle identity (mesh, index) and per-mesh triangle count: rendering/_raycaster.py:16-22,33
 *   - camera: look_at / perspective / viewport / pixel centre: rendering/_core.py:528-548,
 *     rendering/_raster.py:126-129 and :298-299
 *   - barycentric order v0*(1-u-v) + v1*u + v2*v: rendering/_raster.py:46
 *   - Lambert with L = normalize(1,1,1), ambient 0.2: tutorials/lesson09_texture_mapping.py:73
 *   - repeat texture addressing, row = v: rendering/_core.py:94-96
 *
 * Layers:
 *  (1) "twin": the FP32 algorithm of the device code, operation for operation (explicit fmaf,
 *      no contraction: build with -ffp-contract=off), so device output can be compared
 *      bit-for-bit: LBVH build (Morton codes, stable sort, Karras hierarchy, refit), ray
 *      generation, closest/any-hit traversal with two-sided Moller-Trumbore, shading.
 *  (2) "fp64": a brute-force / BVH-culled FP64 closest hit that shares no arithmetic with (1);
 *      it flags rays on which FP32 may legitimately pick another triangle, using a forward error
 *      bound of the FP32 Moller-Trumbore operations: the winner lies within max(edge_eps, bound)
 *      barycentric of an edge, a runner-up is hit within max(tie_eps relative, t bounds) of the
 *      winner (depth tie), or a triangle is missed by less than its bound in that range.
 *  (3) "raster": primary visibility by projection, near-plane clipping, screen-space edge
 *      functions and a depth buffer in FP64 (tw_raster_ids), the reference raster's rule
 *      (rendering/_raster.py:80-112, 172-214, 282-311), a non-ray statement of what a pixel sees.
 *  (4) vertex normals (tw_vertex_normals) for Mesh.compute_normals (rendering/_modeling.py:48-49).
 *
 * Pinning (tests/test_oracle_*.py): Morton vs closed-form KATs; sort vs numpy stable argsort;
 * LBVH invariants (2n-1 nodes, each leaf once, boxes are exact unions) and brute-force equality;
 * FP32 traversal vs FP32 brute force; FP64 closest hit vs closed-form geometry (sphere/plane
 * analytic intersections, axis-aligned triangles) and vs the FP32 twin within 1e-5 relative t;
 * raster vs the FP64 closest hit on unflagged pixels and a projection closed form; normals vs an
 * independent numpy statement and geometric invariants.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TW_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------------------------ */
/* Morton code (30 bits, x in the most significant interleaved position)                       */
/* ------------------------------------------------------------------------------------------ */

TW_API uint32_t tw_expand10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

TW_API uint32_t tw_morton3(uint32_t x, uint32_t y, uint32_t z) {
    return (tw_expand10(x) << 2) | (tw_expand10(y) << 1) | tw_expand10(z);
}

static inline uint32_t quantize10(float c, float lo, float scale) {
    float q = (c - lo) * scale;
    q = fminf(fmaxf(q, 0.0f), 1023.0f);
    return (uint32_t)q;
}

/* ------------------------------------------------------------------------------------------ */
/* Per-triangle preparation (the device gather kernel)                                          */
/* ------------------------------------------------------------------------------------------ */

/* Pads a box by |coordinate| * 2^-14 + 2^-40 per side: wider than the world-space error of an
 * FP32 Moller-Trumbore acceptance, so the box test does not cull a triangle that the FP32 test
 * would accept and the closest hit does not depend on the traversal order. */
static inline float pad_of(float lo, float hi) {
    float m = fmaxf(fabsf(lo), fabsf(hi));
    return m * 0x1p-14f + 0x1p-40f;
}

/* Inputs are flattened per global primitive (mesh-major order):
 *   pos[p*9 + k*3 + c]  vertex k of triangle p (fp32)
 *   nrm[p*9 + ...]      per-vertex normals (ignored where has_nrm[mesh]==0)
 *   uv [p*6 + k*2 + c]  per-vertex texture coordinates (0 where has_uv[mesh]==0)
 *   mesh_of[p]          mesh id
 * Outputs (leaf/slot order after the sort):
 *   keys[n], order[n]   sorted Morton codes and primitive ids
 *   nodes[(n-1)*16]     4 float4 per internal node (layout below)
 *   tris[n*12]          (v0.xyz, prim bits), (e1.xyz, 0), (e2.xyz, 0)
 *   shade[n*16]         (N0.xyz, N1.x), (N1.yz, N2.xy), (N2.z, uv0, uv1.x), (uv1.y, uv2, mesh bits)
 *   bounds[8]           scene AABB lo.xyz,0,hi.xyz,0
 * Node layout: [0]=(b0.lo.x,b0.hi.x,b0.lo.y,b0.hi.y) [1]=(b1.lo.x,b1.hi.x,b1.lo.y,b1.hi.y)
 *              [2]=(b0.lo.z,b0.hi.z,b1.lo.z,b1.hi.z) [3]=(ref0,ref1,0,0) int bits; ref>=0 internal,
 *              ref<0 leaf slot ~ref.
 */

typedef struct { float lo[3], hi[3]; } box3;

static inline float canon(float x) { return x + 0.0f; } /* -0 -> +0 */

static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* Leaves: subtrees of at most LEAF_MAX consecutive slots, referenced as ~(first | (count-1) << 29). */
#define LEAF_MAX 4
static inline int32_t leaf_ref(uint32_t first, uint32_t count) {
    return ~(int32_t)(first | ((count - 1u) << 29));
}
static inline uint32_t leaf_first(int32_t ref) { return (uint32_t)(~ref) & 0x1FFFFFFFu; }
static inline uint32_t leaf_count(int32_t ref) { return ((uint32_t)(~ref) >> 29) + 1u; }

static int delta_k(const uint32_t* k, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a != b) return __builtin_clz(a ^ b);
    return 32 + __builtin_clz((uint32_t)i ^ (uint32_t)j);
}

static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return (x > y) - (x < y);
}

/* Builders (the device's rt_tuning.builder): 0 = LBVH (Karras radix tree over the Morton order),
 * 1 = PLOC (locally ordered clustering over the Morton order, leaves collapsed to <= LEAF_MAX). */
#define TW_BUILD_LBVH 0
#define TW_BUILD_PLOC 1
static int ploc_hierarchy(int64_t n, const box3* leafbox, uint32_t* order, float* nodes, float* bounds);
#define TW_MAX_TREE_DEPTH 62  /* internal levels the device's traversal stacks are sized for */

TW_API int tw_build(int64_t n, const float* pos, const float* nrm, const float* uv,
                    const int32_t* mesh_of, const uint8_t* has_nrm, const uint8_t* has_uv,
                    uint32_t* keys, uint32_t* order, float* nodes, float* nodes4, float* tris,
                    float* shade, float* bounds, int32_t* root_out, int builder) {
    if (n <= 0) {
        for (int c = 0; c < 3; ++c) { bounds[c] = INFINITY; bounds[4 + c] = -INFINITY; }
        bounds[3] = bounds[7] = 0.0f;
        *root_out = 0x7FFFFFFF;
        return 0;
    }
    box3* leafbox = (box3*)malloc(sizeof(box3) * n);
    float* cen = (float*)malloc(sizeof(float) * 3 * n);
    float* tri_u = (float*)malloc(sizeof(float) * 12 * n);
    float* sh_u = (float*)malloc(sizeof(float) * 16 * n);
    if (!leafbox || !cen || !tri_u || !sh_u) return -1;

    /* gather */
    #pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        float v[3][3];
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) v[k][c] = canon(pos[p * 9 + k * 3 + c]);
        box3 b;
        for (int c = 0; c < 3; ++c) {
            float lo = fminf(fminf(v[0][c], v[1][c]), v[2][c]);
            float hi = fmaxf(fmaxf(v[0][c], v[1][c]), v[2][c]);
            cen[p * 3 + c] = (lo + hi) * 0.5f;
            float pd = pad_of(lo, hi);
            b.lo[c] = lo - pd;
            b.hi[c] = hi + pd;
        }
        leafbox[p] = b;
        float e1[3], e2[3];
        for (int c = 0; c < 3; ++c) { e1[c] = v[1][c] - v[0][c]; e2[c] = v[2][c] - v[0][c]; }
        float* t = tri_u + p * 12;
        t[0] = v[0][0]; t[1] = v[0][1]; t[2] = v[0][2]; t[3] = bitsf((uint32_t)p);
        t[4] = e1[0]; t[5] = e1[1]; t[6] = e1[2]; t[7] = 0.0f;
        t[8] = e2[0]; t[9] = e2[1]; t[10] = e2[2]; t[11] = 0.0f;
        int m = mesh_of[p];
        float N[3][3], UV[3][2];
        if (has_nrm[m]) {
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 3; ++c) N[k][c] = nrm[p * 9 + k * 3 + c];
        } else {
            /* geometric normal normalize(e1 x e2) */
            float gx = fmaf(e1[1], e2[2], -(e1[2] * e2[1]));
            float gy = fmaf(e1[2], e2[0], -(e1[0] * e2[2]));
            float gz = fmaf(e1[0], e2[1], -(e1[1] * e2[0]));
            float l2 = fmaf(gx, gx, fmaf(gy, gy, gz * gz));
            float il = l2 > 0.0f ? 1.0f / sqrtf(l2) : 0.0f;
            for (int k = 0; k < 3; ++k) { N[k][0] = gx * il; N[k][1] = gy * il; N[k][2] = gz * il; }
        }
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 2; ++c) UV[k][c] = has_uv[m] ? uv[p * 6 + k * 2 + c] : 0.0f;
        float* s = sh_u + p * 16;
        s[0] = N[0][0]; s[1] = N[0][1]; s[2] = N[0][2]; s[3] = N[1][0];
        s[4] = N[1][1]; s[5] = N[1][2]; s[6] = N[2][0]; s[7] = N[2][1];
        s[8] = N[2][2]; s[9] = UV[0][0]; s[10] = UV[0][1]; s[11] = UV[1][0];
        s[12] = UV[1][1]; s[13] = UV[2][0]; s[14] = UV[2][1]; s[15] = bitsf((uint32_t)m);
    }

    /* centroid bounds and Morton codes */
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t p = 0; p < n; ++p)
        for (int c = 0; c < 3; ++c) {
            clo[c] = fminf(clo[c], cen[p * 3 + c]);
            chi[c] = fmaxf(chi[c], cen[p * 3 + c]);
        }
    float sc[3];
    for (int c = 0; c < 3; ++c) {
        float ext = chi[c] - clo[c];
        sc[c] = ext > 0.0f ? 1024.0f / ext : 0.0f;
    }
    uint64_t* kv = (uint64_t*)malloc(sizeof(uint64_t) * n);
    if (!kv) return -1;
    #pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        uint32_t qx = quantize10(cen[p * 3 + 0], clo[0], sc[0]);
        uint32_t qy = quantize10(cen[p * 3 + 1], clo[1], sc[1]);
        uint32_t qz = quantize10(cen[p * 3 + 2], clo[2], sc[2]);
        kv[p] = ((uint64_t)tw_morton3(qx, qy, qz) << 32) | (uint64_t)p;
    }
    /* stable sort by key == sort by (key, primitive id) */
    qsort(kv, (size_t)n, sizeof(uint64_t), cmp_u64);
    for (int64_t i = 0; i < n; ++i) { keys[i] = (uint32_t)(kv[i] >> 32); order[i] = (uint32_t)kv[i]; }
    free(kv);

    if (n == 1) {
        memcpy(tris, tri_u + (int64_t)order[0] * 12, 48);
        memcpy(shade, sh_u + (int64_t)order[0] * 16, 64);
        box3 b = leafbox[order[0]];
        for (int c = 0; c < 3; ++c) { bounds[c] = b.lo[c]; bounds[4 + c] = b.hi[c]; }
        bounds[3] = bounds[7] = 0.0f;
        *root_out = ~0;
        free(leafbox); free(cen); free(tri_u); free(sh_u);
        return 0;
    }

    int lbvh = 1;
    if (builder == TW_BUILD_PLOC) {
        const int rc = ploc_hierarchy(n, leafbox, order, nodes, bounds);
        if (rc < 0) return -1;
        lbvh = rc == 1;  /* too deep: the LBVH of the same keys, as the device does */
    }
    if (lbvh) {
    /* Karras 2012 binary radix tree over the 62-bit extended keys (code, slot) */
    int32_t* left = (int32_t*)malloc(sizeof(int32_t) * (n - 1));
    int32_t* right = (int32_t*)malloc(sizeof(int32_t) * (n - 1));
    int32_t* tleft = (int32_t*)malloc(sizeof(int32_t) * (n - 1));
    int32_t* tright = (int32_t*)malloc(sizeof(int32_t) * (n - 1));
    if (!left || !right || !tleft || !tright) return -1;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n - 1; ++i) {
        int d = (delta_k(keys, n, i, i + 1) - delta_k(keys, n, i, i - 1)) >= 0 ? 1 : -1;
        int dmin = delta_k(keys, n, i, i - d);
        int64_t lmax = 2;
        while (delta_k(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
        int64_t l = 0;
        for (int64_t t = lmax / 2; t >= 1; t /= 2)
            if (delta_k(keys, n, i, i + (l + t) * d) > dmin) l += t;
        int64_t j = i + l * d;
        int dnode = delta_k(keys, n, i, j);
        int64_t s = 0;
        int64_t t = l;
        for (;;) {
            t = (t + 1) / 2;
            if (delta_k(keys, n, i, i + (s + t) * d) > dnode) s += t;
            if (t <= 1) break;
        }
        int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
        int64_t lo_ = i < j ? i : j, hi_ = i < j ? j : i;
        left[i] = (lo_ == gamma) ? ~(int32_t)gamma : (int32_t)gamma;
        right[i] = (hi_ == gamma + 1) ? ~(int32_t)(gamma + 1) : (int32_t)(gamma + 1);
        int64_t nl = gamma - lo_ + 1, nr = hi_ - gamma;
        tleft[i] = nl <= LEAF_MAX ? leaf_ref((uint32_t)lo_, (uint32_t)nl) : (int32_t)gamma;
        tright[i] = nr <= LEAF_MAX ? leaf_ref((uint32_t)(gamma + 1), (uint32_t)nr) : (int32_t)(gamma + 1);
    }

    /* refit: post-order over the tree, child boxes stored in the parent */
    box3* nb = (box3*)malloc(sizeof(box3) * (n - 1));
    int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * 2 * (n + 1));
    uint8_t* state = (uint8_t*)calloc((size_t)(n - 1), 1);
    if (!nb || !stack || !state) return -1;
    int64_t sp = 0;
    stack[sp++] = 0;
    while (sp) {
        int32_t x = stack[sp - 1];
        if (state[x] == 0) {
            state[x] = 1;
            if (right[x] >= 0) stack[sp++] = right[x];
            if (left[x] >= 0) stack[sp++] = left[x];
            continue;
        }
        --sp;
        box3 b0 = left[x] >= 0 ? nb[left[x]] : leafbox[order[~left[x]]];
        box3 b1 = right[x] >= 0 ? nb[right[x]] : leafbox[order[~right[x]]];
        float* nd = nodes + (int64_t)x * 16;
        nd[0] = b0.lo[0]; nd[1] = b0.hi[0]; nd[2] = b0.lo[1]; nd[3] = b0.hi[1];
        nd[4] = b1.lo[0]; nd[5] = b1.hi[0]; nd[6] = b1.lo[1]; nd[7] = b1.hi[1];
        nd[8] = b0.lo[2]; nd[9] = b0.hi[2]; nd[10] = b1.lo[2]; nd[11] = b1.hi[2];
        nd[12] = bitsf((uint32_t)tleft[x]); nd[13] = bitsf((uint32_t)tright[x]);
        nd[14] = 0.0f; nd[15] = 0.0f;
        box3 u;
        for (int c = 0; c < 3; ++c) {
            u.lo[c] = fminf(b0.lo[c], b1.lo[c]);
            u.hi[c] = fmaxf(b0.hi[c], b1.hi[c]);
        }
        nb[x] = u;
    }
    for (int c = 0; c < 3; ++c) { bounds[c] = nb[0].lo[c]; bounds[4 + c] = nb[0].hi[c]; }
    bounds[3] = bounds[7] = 0.0f;
    free(stack); free(state); free(nb); free(left); free(right); free(tleft); free(tright);
    }
    /* leaf records in slot order */
    #pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < n; ++s) {
        uint32_t p = order[s];
        memcpy(tris + s * 12, tri_u + (int64_t)p * 12, 48);
        memcpy(shade + s * 16, sh_u + (int64_t)p * 16, 64);
    }
    *root_out = n <= LEAF_MAX ? leaf_ref(0, (uint32_t)n) : 0;
    /* 4-wide records: grandchildren through internal children, leaf children directly */
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n - 1; ++i) {
        float lo[3][4], hi[3][4];
        int32_t ref[4];
        int k = 0;
        const float* nd = nodes + i * 16;
        int32_t c[2] = {(int32_t)fbits(nd[12]), (int32_t)fbits(nd[13])};
        float cb[2][6] = {{nd[0], nd[2], nd[8], nd[1], nd[3], nd[9]},
                          {nd[4], nd[6], nd[10], nd[5], nd[7], nd[11]}};
        for (int e = 0; e < 2; ++e) {
            if (c[e] >= 0) {
                const float* cd = nodes + (int64_t)c[e] * 16;
                lo[0][k] = cd[0]; hi[0][k] = cd[1]; lo[1][k] = cd[2]; hi[1][k] = cd[3];
                lo[2][k] = cd[8]; hi[2][k] = cd[9]; ref[k] = (int32_t)fbits(cd[12]); ++k;
                lo[0][k] = cd[4]; hi[0][k] = cd[5]; lo[1][k] = cd[6]; hi[1][k] = cd[7];
                lo[2][k] = cd[10]; hi[2][k] = cd[11]; ref[k] = (int32_t)fbits(cd[13]); ++k;
            } else {
                for (int a = 0; a < 3; ++a) { lo[a][k] = cb[e][a]; hi[a][k] = cb[e][3 + a]; }
                ref[k] = c[e]; ++k;
            }
        }
        for (; k < 4; ++k) {
            for (int a = 0; a < 3; ++a) { lo[a][k] = INFINITY; hi[a][k] = -INFINITY; }
            ref[k] = 0x7FFFFFFF;
        }
        float* o = nodes4 + i * 32;
        for (int a = 0; a < 3; ++a)
            for (int q = 0; q < 4; ++q) { o[8 * a + q] = lo[a][q]; o[8 * a + 4 + q] = hi[a][q]; }
        for (int q = 0; q < 4; ++q) o[24 + q] = bitsf((uint32_t)ref[q]);
        for (int q = 0; q < 4; ++q) o[28 + q] = 0.0f;
    }
    free(leafbox); free(cen); free(tri_u); free(sh_u);
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* PLOC: parallel locally-ordered clustering (Meister & Bittner 2018), the device's iteration   */
/* ------------------------------------------------------------------------------------------ */
/* Clusters start as the triangles in Morton order. Each iteration: every cluster i finds its
 * nearest neighbour NN(i) among positions [i-R, i+R] (smallest surface area of the union box; ties
 * to the smaller position); mutual pairs (NN(NN(i)) == i) merge into a new node at the lower
 * position, the other position drops out; the survivors keep their order. An iteration without a
 * merge (only possible with NaN boxes) merges positions 0 and 1. The k-th node created (counting
 * over iterations in position order) gets id n-2-k, so the root is 0 and children have larger
 * ids than their parents. Then subtrees of <= LEAF_MAX triangles whose parent is larger become
 * leaves ("leaf roots"): the leaf roots take consecutive slot ranges in the order of the smallest
 * Morton position of their triangles, and the triangles of a leaf root take its range in
 * left-to-right order. */
#define PLOC_RADIUS 16

static inline float union_area(const box3* a, const box3* b) {
    float dx = fmaxf(a->hi[0], b->hi[0]) - fminf(a->lo[0], b->lo[0]);
    float dy = fmaxf(a->hi[1], b->hi[1]) - fminf(a->lo[1], b->lo[1]);
    float dz = fmaxf(a->hi[2], b->hi[2]) - fminf(a->lo[2], b->lo[2]);
    return dx * dy + dy * dz + dz * dx;
}

typedef struct {
    int64_t n;
    const int32_t *left, *right, *size;
    const uint32_t* pos_of;
} ploc_tree;

static inline int32_t ploc_size(const ploc_tree* t, int32_t c) { return c >= 0 ? t->size[c] : 1; }

static uint32_t ploc_minpos(const ploc_tree* t, int32_t c) {
    if (c < 0) return t->pos_of[~c];
    uint32_t a = ploc_minpos(t, t->left[c]), b = ploc_minpos(t, t->right[c]);
    return a < b ? a : b;
}

static void ploc_slots(const ploc_tree* t, int32_t c, uint32_t* next, uint32_t* order,
                       uint32_t* slot_of, uint32_t* fslot) {
    if (c < 0) {
        order[*next] = (uint32_t)~c;
        slot_of[~c] = *next;
        ++*next;
        return;
    }
    fslot[c] = *next;
    ploc_slots(t, t->left[c], next, order, slot_of, fslot);
    ploc_slots(t, t->right[c], next, order, slot_of, fslot);
}

static int ploc_hierarchy(int64_t n, const box3* leafbox, uint32_t* order, float* nodes, float* bounds) {
    int32_t* C = (int32_t*)malloc(sizeof(int32_t) * n);
    int32_t* C2 = (int32_t*)malloc(sizeof(int32_t) * n);
    int64_t* NN = (int64_t*)malloc(sizeof(int64_t) * n);
    box3* B = (box3*)malloc(sizeof(box3) * n);
    box3* B2 = (box3*)malloc(sizeof(box3) * n);
    int32_t* left = (int32_t*)malloc(sizeof(int32_t) * (n - 1));
    int32_t* right = (int32_t*)malloc(sizeof(int32_t) * (n - 1));
    int32_t* size = (int32_t*)malloc(sizeof(int32_t) * (n - 1));
    int32_t* parent = (int32_t*)malloc(sizeof(int32_t) * (2 * n - 1)); /* internal ids, then n-1+p */
    uint32_t* pos_of = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* morder = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* slot_of = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* fslot = (uint32_t*)malloc(sizeof(uint32_t) * (n - 1));
    if (!C || !C2 || !NN || !B || !B2 || !left || !right || !size || !parent || !pos_of || !morder ||
        !slot_of || !fslot)
        return -1;
    for (int64_t i = 0; i < n; ++i) {
        morder[i] = order[i];
        pos_of[order[i]] = (uint32_t)i;
        C[i] = ~(int32_t)order[i];
        B[i] = leafbox[order[i]];
    }
    int64_t m = n, created = 0;
    while (m > 1) {
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < m; ++i) {
            int64_t lo = i - PLOC_RADIUS < 0 ? 0 : i - PLOC_RADIUS;
            int64_t hi = i + PLOC_RADIUS > m - 1 ? m - 1 : i + PLOC_RADIUS;
            int64_t bj = lo == i ? i + 1 : lo;
            float best = union_area(&B[i], &B[bj]);
            for (int64_t j = bj + 1; j <= hi; ++j) {
                if (j == i) continue;
                float a = union_area(&B[i], &B[j]);
                if (a < best) { best = a; bj = j; }
            }
            NN[i] = bj;
        }
        int any = 0;
        for (int64_t i = 0; i < m && !any; ++i) any = NN[NN[i]] == i;
        if (!any) { NN[0] = 1; NN[1] = 0; }
        int64_t k = 0;
        for (int64_t i = 0; i < m; ++i) {
            int64_t j = NN[i];
            if (NN[j] == i) {
                if (j < i) continue;
                int32_t id = (int32_t)(n - 2 - created);
                ++created;
                left[id] = C[i];
                right[id] = C[j];
                size[id] = (C[i] >= 0 ? size[C[i]] : 1) + (C[j] >= 0 ? size[C[j]] : 1);
                parent[C[i] >= 0 ? C[i] : n - 1 + ~C[i]] = id;
                parent[C[j] >= 0 ? C[j] : n - 1 + ~C[j]] = id;
                const box3* a = &B[i];
                const box3* b = &B[j];
                float* nd = nodes + (int64_t)id * 16;
                nd[0] = a->lo[0]; nd[1] = a->hi[0]; nd[2] = a->lo[1]; nd[3] = a->hi[1];
                nd[4] = b->lo[0]; nd[5] = b->hi[0]; nd[6] = b->lo[1]; nd[7] = b->hi[1];
                nd[8] = a->lo[2]; nd[9] = a->hi[2]; nd[10] = b->lo[2]; nd[11] = b->hi[2];
                nd[14] = 0.0f; nd[15] = 0.0f;
                box3 u;
                for (int c = 0; c < 3; ++c) { u.lo[c] = fminf(a->lo[c], b->lo[c]); u.hi[c] = fmaxf(a->hi[c], b->hi[c]); }
                C2[k] = id;
                B2[k] = u;
                ++k;
            } else {
                C2[k] = C[i];
                B2[k] = B[i];
                ++k;
            }
        }
        int32_t* tc = C; C = C2; C2 = tc;
        box3* tb = B; B = B2; B2 = tb;
        m = k;
    }
    parent[0] = -1;
    /* the device's depth guard: a tree with a triangle under more than TW_MAX_TREE_DEPTH internal
       nodes is not used (the caller builds the LBVH instead); `order` is still the Morton order */
    for (int64_t p = 0; p < n; ++p) {
        int32_t x = parent[n - 1 + p];
        int d = 0;
        while (x >= 0 && d <= TW_MAX_TREE_DEPTH) { ++d; x = parent[x]; }
        if (d > TW_MAX_TREE_DEPTH) {
            free(C); free(C2); free(NN); free(B); free(B2); free(left); free(right); free(size);
            free(parent); free(pos_of); free(morder); free(slot_of); free(fslot);
            return 1;
        }
    }
    for (int c = 0; c < 3; ++c) { bounds[c] = B[0].lo[c]; bounds[4 + c] = B[0].hi[c]; }
    bounds[3] = bounds[7] = 0.0f;
    /* leaf roots in the order of their smallest Morton position; slots left to right */
    ploc_tree t = {n, left, right, size, pos_of};
    uint32_t next = 0;
    for (int64_t i = 0; i < n; ++i) {
        int32_t x = ~(int32_t)morder[i];
        for (;;) {
            int32_t p = parent[x >= 0 ? x : n - 1 + ~x];
            if (p < 0 || size[p] > LEAF_MAX) break;
            x = p;
        }
        if (ploc_minpos(&t, x) != (uint32_t)i) continue;
        ploc_slots(&t, x, &next, order, slot_of, fslot);
    }
    /* traversal refs: a child of more than LEAF_MAX triangles by id, else the leaf of its slots */
    for (int64_t x = 0; x < n - 1; ++x) {
        int32_t ch[2] = {left[x], right[x]};
        for (int e = 0; e < 2; ++e) {
            int32_t c = ch[e], r;
            if (c < 0) r = leaf_ref(slot_of[~c], 1);
            else if (size[c] > LEAF_MAX) r = c;
            else r = leaf_ref(fslot[c], (uint32_t)size[c]);
            nodes[x * 16 + 12 + e] = bitsf((uint32_t)r);
        }
    }
    free(C); free(C2); free(NN); free(B); free(B2); free(left); free(right); free(size);
    free(parent); free(pos_of); free(morder); free(slot_of); free(fslot);
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* FP32 traversal (the device traversal, operation for operation)                             */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    float ox, oy, oz, dx, dy, dz, tmin, tmax;
} ray32;

static inline float safe_inv(float d) {
    float dd = fabsf(d) < 0x1p-80f ? copysignf(0x1p-80f, d) : d;
    return 1.0f / dd;
}

/* 1 + 2*gamma(3), gamma(n) = n*eps/(1-n*eps), eps = 2^-24, rounded up (Ize 2013). */
#define TW_TFAR_SCALE 1.0000004f

/* Two-sided Moller-Trumbore; the barycentric tests compare the numerators with the determinant
 * after multiplying both by sign(det) (exact), the division happens only for a hit. */
static inline int mt32(const float* tr, const ray32* r, float* t, float* u, float* v) {
    float e1x = tr[4], e1y = tr[5], e1z = tr[6];
    float e2x = tr[8], e2y = tr[9], e2z = tr[10];
    float px = fmaf(r->dy, e2z, -(r->dz * e2y));
    float py = fmaf(r->dz, e2x, -(r->dx * e2z));
    float pz = fmaf(r->dx, e2y, -(r->dy * e2x));
    float det = fmaf(e1x, px, fmaf(e1y, py, e1z * pz));
    if (det == 0.0f) return 0;
    float sg = det > 0.0f ? 1.0f : -1.0f, ad = det * sg;
    float sx = r->ox - tr[0], sy = r->oy - tr[1], sz = r->oz - tr[2];
    float un = fmaf(sx, px, fmaf(sy, py, sz * pz));
    float us = un * sg;
    if (us < 0.0f || us > ad) return 0;
    float qx = fmaf(sy, e1z, -(sz * e1y));
    float qy = fmaf(sz, e1x, -(sx * e1z));
    float qz = fmaf(sx, e1y, -(sy * e1x));
    float vn = fmaf(r->dx, qx, fmaf(r->dy, qy, r->dz * qz));
    float vs = vn * sg;
    if (vs < 0.0f || us + vs > ad) return 0;
    float inv = 1.0f / det;
    *t = fmaf(e2x, qx, fmaf(e2y, qy, e2z * qz)) * inv;
    *u = un * inv;
    *v = vn * inv;
    return 1;
}

typedef struct {
    float t, u, v;
    int32_t prim, slot;
    int32_t node_visits, tri_tests;
} hit32;

/* any != 0: stop at the first accepted hit (occlusion); skip_slot is never tested. */
static void traverse32(const float* nodes, const float* tris, int32_t root, const ray32* r,
                       int any, int32_t skip_slot, hit32* h) {
    h->t = r->tmax; h->u = 0.0f; h->v = 0.0f; h->prim = -1; h->slot = -1;
    h->node_visits = 0; h->tri_tests = 0;
    if (root == 0x7FFFFFFF) return;
    float idx = safe_inv(r->dx), idy = safe_inv(r->dy), idz = safe_inv(r->dz);
    float oxx = r->ox * idx, oyy = r->oy * idy, ozz = r->oz * idz;
    int32_t stack[64];
    int sp = 0;
    int32_t cur = root;
    for (;;) {
        if (cur >= 0) {
            const float* nd = nodes + (int64_t)cur * 16;
            h->node_visits++;
            float a0 = fmaf(nd[0], idx, -oxx), a1 = fmaf(nd[1], idx, -oxx);
            float a2 = fmaf(nd[2], idy, -oyy), a3 = fmaf(nd[3], idy, -oyy);
            float a4 = fmaf(nd[8], idz, -ozz), a5 = fmaf(nd[9], idz, -ozz);
            float b0 = fmaf(nd[4], idx, -oxx), b1 = fmaf(nd[5], idx, -oxx);
            float b2 = fmaf(nd[6], idy, -oyy), b3 = fmaf(nd[7], idy, -oyy);
            float b4 = fmaf(nd[10], idz, -ozz), b5 = fmaf(nd[11], idz, -ozz);
            float tn0 = fmaxf(fmaxf(fminf(a0, a1), fminf(a2, a3)), fmaxf(fminf(a4, a5), r->tmin));
            float tf0 = fminf(fminf(fmaxf(a0, a1), fmaxf(a2, a3)), fmaxf(a4, a5)) * TW_TFAR_SCALE;
            float tn1 = fmaxf(fmaxf(fminf(b0, b1), fminf(b2, b3)), fmaxf(fminf(b4, b5), r->tmin));
            float tf1 = fminf(fminf(fmaxf(b0, b1), fmaxf(b2, b3)), fmaxf(b4, b5)) * TW_TFAR_SCALE;
            int hit0 = tn0 <= fminf(tf0, h->t);
            int hit1 = tn1 <= fminf(tf1, h->t);
            int32_t r0 = (int32_t)fbits(nd[12]), r1 = (int32_t)fbits(nd[13]);
            if (hit0 && hit1) {
                int32_t nearc = r0, farc = r1;
                if (tn1 < tn0) { nearc = r1; farc = r0; }
                if (sp < 64) stack[sp++] = farc;
                cur = nearc;
                continue;
            }
            if (hit0) { cur = r0; continue; }
            if (hit1) { cur = r1; continue; }
        } else {
            uint32_t first = leaf_first(cur), count = leaf_count(cur);
            for (uint32_t k = 0; k < count; ++k) {
                int32_t slot = (int32_t)(first + k);
                if (slot == skip_slot) continue;
                const float* tr = tris + (int64_t)slot * 12;
                float t, u, v;
                h->tri_tests++;
                if (mt32(tr, r, &t, &u, &v) && t > r->tmin) {
                    int32_t prim = (int32_t)fbits(tr[3]);
                    if (t < h->t || (t == h->t && prim < h->prim)) {
                        h->t = t; h->u = u; h->v = v; h->prim = prim; h->slot = slot;
                        if (any) return;
                    }
                }
            }
        }
        if (sp == 0) break;
        cur = stack[--sp];
    }
}

/* FP64 re-evaluation of the winner's (t, u, v) from its stored FP32 record (the device's
 * refine_hit, operation for operation). */
static void refine64(const float* tris, const ray32* r, hit32* h) {
    const float* tr = tris + (int64_t)h->slot * 12;
    double dx = r->dx, dy = r->dy, dz = r->dz;
    double px = fma(dy, (double)tr[10], -(dz * (double)tr[9]));
    double py = fma(dz, (double)tr[8], -(dx * (double)tr[10]));
    double pz = fma(dx, (double)tr[9], -(dy * (double)tr[8]));
    double det = fma((double)tr[4], px, fma((double)tr[5], py, (double)tr[6] * pz));
    if (det == 0.0) return;
    double sx = (double)r->ox - (double)tr[0], sy = (double)r->oy - (double)tr[1],
           sz = (double)r->oz - (double)tr[2];
    double qx = fma(sy, (double)tr[6], -(sz * (double)tr[5]));
    double qy = fma(sz, (double)tr[4], -(sx * (double)tr[6]));
    double qz = fma(sx, (double)tr[5], -(sy * (double)tr[4]));
    double un = fma(sx, px, fma(sy, py, sz * pz));
    double vn = fma(dx, qx, fma(dy, qy, dz * qz));
    double tn = fma((double)tr[8], qx, fma((double)tr[9], qy, (double)tr[10] * qz));
    h->t = (float)(tn / det);
    h->u = (float)(un / det);
    h->v = (float)(vn / det);
}

static inline void load_ray(const float* ox, const float* oy, const float* oz, const float* dx,
                            const float* dy, const float* dz, const float* tmin,
                            const float* tmax, float tmin_d, float tmax_d, int64_t i, ray32* r) {
    r->ox = ox[i]; r->oy = oy[i]; r->oz = oz[i];
    r->dx = dx[i]; r->dy = dy[i]; r->dz = dz[i];
    r->tmin = tmin ? tmin[i] : tmin_d;
    r->tmax = tmax ? tmax[i] : tmax_d;
}

/* Closest hit of n SoA rays. Outputs may be NULL. counters (may be NULL): per-ray node visits
 * and triangle tests, 2 int32 per ray. */
TW_API void tw_trace_closest(const float* nodes, const float* tris, int32_t root, int64_t n,
                             const float* ox, const float* oy, const float* oz, const float* dx,
                             const float* dy, const float* dz, const float* tmin,
                             const float* tmax, float tmin_d, float tmax_d, float* out_t,
                             float* out_u, float* out_v, int32_t* out_prim, int32_t* out_slot,
                             int32_t* counters) {
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        ray32 r;
        load_ray(ox, oy, oz, dx, dy, dz, tmin, tmax, tmin_d, tmax_d, i, &r);
        hit32 h;
        traverse32(nodes, tris, root, &r, 0, -1, &h);
        if (h.prim >= 0) refine64(tris, &r, &h);
        if (h.prim < 0) { h.t = INFINITY; h.u = 0.0f; h.v = 0.0f; }
        if (out_t) out_t[i] = h.t;
        if (out_u) out_u[i] = h.u;
        if (out_v) out_v[i] = h.v;
        if (out_prim) out_prim[i] = h.prim;
        if (out_slot) out_slot[i] = h.slot;
        if (counters) { counters[2 * i] = h.node_visits; counters[2 * i + 1] = h.tri_tests; }
    }
}

TW_API void tw_trace_any(const float* nodes, const float* tris, int32_t root, int64_t n,
                         const float* ox, const float* oy, const float* oz, const float* dx,
                         const float* dy, const float* dz, const float* tmin, const float* tmax,
                         float tmin_d, float tmax_d, uint8_t* occluded) {
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        ray32 r;
        load_ray(ox, oy, oz, dx, dy, dz, tmin, tmax, tmin_d, tmax_d, i, &r);
        hit32 h;
        traverse32(nodes, tris, root, &r, 1, -1, &h);
        occluded[i] = h.prim >= 0;
    }
}

/* FP32 brute force over every slot with the same Moller-Trumbore and tie rule. */
TW_API void tw_brute_closest32(const float* tris, int64_t n_tris, int64_t n, const float* ox,
                               const float* oy, const float* oz, const float* dx, const float* dy,
                               const float* dz, float tmin_d, float tmax_d, float* out_t,
                               int32_t* out_prim) {
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i) {
        ray32 r = {ox[i], oy[i], oz[i], dx[i], dy[i], dz[i], tmin_d, tmax_d};
        float bt = tmax_d;
        int32_t bp = -1;
        for (int64_t s = 0; s < n_tris; ++s) {
            const float* tr = tris + s * 12;
            float t, u, v;
            if (mt32(tr, &r, &t, &u, &v) && t > r.tmin) {
                int32_t prim = (int32_t)fbits(tr[3]);
                if (t < bt || (t == bt && prim < bp)) { bt = t; bp = prim; }
            }
        }
        out_t[i] = bp >= 0 ? bt : INFINITY;
        out_prim[i] = bp;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Ray generation and shading (FP32, the device formulas)                                       */
/* ------------------------------------------------------------------------------------------ */

/* cam: eye[3], right[3], up[3], forward[3] (see rt_camera in include/rendertoy_rt.h). */
static inline void gen_ray(const float* cam, uint32_t W, uint32_t H, uint32_t col, uint32_t row,
                           float* d) {
    float nx = ((float)col + 0.5f) / (float)W * 2.0f - 1.0f;
    float ny = 1.0f - ((float)row + 0.5f) / (float)H * 2.0f;
    float x = fmaf(nx, cam[3], fmaf(ny, cam[6], cam[9]));
    float y = fmaf(nx, cam[4], fmaf(ny, cam[7], cam[10]));
    float z = fmaf(nx, cam[5], fmaf(ny, cam[8], cam[11]));
    float il = 1.0f / sqrtf(fmaf(x, x, fmaf(y, y, z * z)));
    d[0] = x * il; d[1] = y * il; d[2] = z * il;
}

TW_API void tw_generate_rays(const float* cam, uint32_t W, uint32_t H, float* ox, float* oy,
                             float* oz, float* dx, float* dy, float* dz) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)W * H; ++i) {
        float d[3];
        gen_ray(cam, W, H, (uint32_t)(i % W), (uint32_t)(i / W), d);
        ox[i] = cam[0]; oy[i] = cam[1]; oz[i] = cam[2];
        dx[i] = d[0]; dy[i] = d[1]; dz[i] = d[2];
    }
}

static inline int wrap_i(int i, int n) { int r = i % n; return r < 0 ? r + n : r; }

/* RGBA8 texture fetch: bilinear (OpenCL 1.2 linear filter, repeat addressing) or point
 * sampling as rendering/_core.py:94-96 (texel (int)(wrap(v)*H), (int)(wrap(u)*W)). */
static inline void tex_sample(const uint8_t* tex, int W, int H, float u, float v, int bilinear,
                              float* rgb) {
    float su = u - floorf(u), sv = v - floorf(v);
    if (!bilinear) {
        int i = (int)(su * (float)W), j = (int)(sv * (float)H);
        if (i > W - 1) i = W - 1;
        if (j > H - 1) j = H - 1;
        const uint8_t* t = tex + ((int64_t)j * W + i) * 4;
        for (int c = 0; c < 3; ++c) rgb[c] = (float)t[c] * (1.0f / 255.0f);
        return;
    }
    float x = su * (float)W - 0.5f, y = sv * (float)H - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float a = x - fx, b = y - fy;
    int i0 = wrap_i((int)fx, W), j0 = wrap_i((int)fy, H);
    int i1 = wrap_i((int)fx + 1, W), j1 = wrap_i((int)fy + 1, H);
    const uint8_t* t00 = tex + ((int64_t)j0 * W + i0) * 4;
    const uint8_t* t10 = tex + ((int64_t)j0 * W + i1) * 4;
    const uint8_t* t01 = tex + ((int64_t)j1 * W + i0) * 4;
    const uint8_t* t11 = tex + ((int64_t)j1 * W + i1) * 4;
    for (int c = 0; c < 3; ++c) {
        float top = fmaf(a, (float)t10[c] - (float)t00[c], (float)t00[c]);
        float bot = fmaf(a, (float)t11[c] - (float)t01[c], (float)t01[c]);
        rgb[c] = fmaf(b, bot - top, top) * (1.0f / 255.0f);
    }
}

/* Texture fetch of the device shading kernel, exposed for known-answer tests. */
TW_API void tw_tex_sample(const uint8_t* tex, int32_t W, int32_t H, int64_t n, const float* u,
                          const float* v, int32_t bilinear, float* rgb) {
    for (int64_t i = 0; i < n; ++i) tex_sample(tex, W, H, u[i], v[i], bilinear, rgb + 3 * i);
}

/* ------------------------------------------------------------------------------------------ */
/* Wave-coherent traversal (the device traverse_packet), one 64-lane wave emulated lane by lane */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    int64_t node_visits, leaf_visits, tri_fetches;  /* per wave: 4-wide node records, leaves and
                                                        triangle records fetched */
} wavecount;

/* The device's packet traversal (traverse_packet), lane-exact: node steps test the 4 children of
 * the current node for the lanes in the mask; hit leaf children are tested in the same step (per
 * lane in the child's mask, in child order); internal children are ordered by the lowest active
 * lane's entry distances (sorting network on (key, k)), the nearest is visited next, the others are
 * pushed; with no internal child hit the next node is popped before the leaf tests (any-hit:
 * entries whose lanes have all found an occluder are skipped, re-checked at the next step). */
static void packet32(const float* nodes4, const float* tris, int32_t root, const ray32* r,
                     const int* want, const int32_t* skip, int any, hit32* h, wavecount* wc) {
    uint64_t m = 0;
    float idx[64], idy[64], idz[64], oxx[64], oyy[64], ozz[64];
    for (int l = 0; l < 64; ++l) {
        h[l].t = r[l].tmax; h[l].u = 0.0f; h[l].v = 0.0f; h[l].prim = -1; h[l].slot = -1;
        h[l].node_visits = 0; h[l].tri_tests = 0;
        if (want[l]) m |= 1ull << l;
        idx[l] = safe_inv(r[l].dx); idy[l] = safe_inv(r[l].dy); idz[l] = safe_inv(r[l].dz);
        oxx[l] = r[l].ox * idx[l]; oyy[l] = r[l].oy * idy[l]; ozz[l] = r[l].oz * idz[l];
    }
    if (root == 0x7FFFFFFF || m == 0) return;
    if (root < 0) {  /* single-leaf tree */
        uint32_t first = leaf_first(root), count = leaf_count(root);
        if (wc) { wc->leaf_visits++; wc->tri_fetches += count; }
        for (uint32_t k = 0; k < count; ++k) {
            int32_t slot = (int32_t)(first + k);
            const float* tr = tris + (int64_t)slot * 12;
            for (int l = 0; l < 64; ++l) {
                if (!((m >> l) & 1) || slot == skip[l]) continue;
                float t, u, v;
                h[l].tri_tests++;
                if (mt32(tr, &r[l], &t, &u, &v) && t > r[l].tmin) {
                    int32_t prim = (int32_t)fbits(tr[3]);
                    if (t < h[l].t || (t == h[l].t && prim < h[l].prim)) {
                        h[l].t = t; h[l].u = u; h[l].v = v; h[l].prim = prim; h[l].slot = slot;
                    }
                }
            }
        }
        return;
    }
    uint64_t alive = m;
    int32_t stk_node[128];
    uint64_t stk_m[128];
    int sp = 0;
    int32_t cur = root;
    for (;;) {
        if (any) {
            m &= alive;
            while (m == 0) {
                if (sp == 0) return;
                --sp;
                cur = stk_node[sp];
                m = stk_m[sp] & alive;
            }
        }
        const float* nd = nodes4 + (int64_t)cur * 32;
        if (wc) wc->node_visits++;
        int rep = __builtin_ctzll(m);
        float key[4];
        uint64_t mk[4];
        int32_t ref[4];
        for (int k = 0; k < 4; ++k) {
            ref[k] = (int32_t)fbits(nd[24 + k]);
            mk[k] = 0;
            key[k] = INFINITY;
        }
        for (int l = 0; l < 64; ++l) {
            if (!((m >> l) & 1)) continue;
            h[l].node_visits++;
            for (int k = 0; k < 4; ++k) {
                float x0 = fmaf(nd[k], idx[l], -oxx[l]), x1 = fmaf(nd[4 + k], idx[l], -oxx[l]);
                float y0 = fmaf(nd[8 + k], idy[l], -oyy[l]), y1 = fmaf(nd[12 + k], idy[l], -oyy[l]);
                float z0 = fmaf(nd[16 + k], idz[l], -ozz[l]), z1 = fmaf(nd[20 + k], idz[l], -ozz[l]);
                float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), r[l].tmin));
                float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1)) * TW_TFAR_SCALE;
                int hk = ref[k] != 0x7FFFFFFF && tn <= fminf(tf, h[l].t);
                if (hk) mk[k] |= 1ull << l;
                if (l == rep) key[k] = hk ? tn : INFINITY;
            }
        }
        /* internal children in (key, k) order by the device's sorting network */
        int ord[4] = {0, 1, 2, 3};
        const int net[5][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}};
        for (int e = 0; e < 5; ++e) {
            int a = net[e][0], b = net[e][1];
            int ka = ord[a], kb = ord[b];
            if (key[kb] < key[ka] || (key[kb] == key[ka] && kb < ka)) { ord[a] = kb; ord[b] = ka; }
        }
        int32_t next = -1;
        uint64_t next_m = 0;
        for (int q = 3; q >= 0; --q) {
            int k = ord[q];
            uint64_t im = ref[k] >= 0 ? mk[k] : 0;
            if (!im) continue;
            if (next_m && sp < 128) { stk_node[sp] = next; stk_m[sp] = next_m; ++sp; }
            next = ref[k];
            next_m = im;
        }
        int more = next_m != 0;
        while (!more && sp > 0) {  /* pop before the leaf tests, as the device does */
            --sp;
            next = stk_node[sp];
            next_m = stk_m[sp];
            if (any) next_m &= alive;
            more = next_m != 0;
        }
        /* the hit leaf children, in child order */
        int tested = 0;
        for (int k = 0; k < 4; ++k) {
            if (ref[k] >= 0 || !mk[k]) continue;
            tested = 1;
            uint32_t first = leaf_first(ref[k]), count = leaf_count(ref[k]);
            if (wc) { wc->leaf_visits++; wc->tri_fetches += count; }
            for (uint32_t t = 0; t < count; ++t) {
                int32_t slot = (int32_t)(first + t);
                const float* tr = tris + (int64_t)slot * 12;
                for (int l = 0; l < 64; ++l) {
                    if (!((mk[k] >> l) & 1) || slot == skip[l]) continue;
                    float tt, u, v;
                    h[l].tri_tests++;
                    if (mt32(tr, &r[l], &tt, &u, &v) && tt > r[l].tmin) {
                        int32_t prim = (int32_t)fbits(tr[3]);
                        if (tt < h[l].t || (tt == h[l].t && prim < h[l].prim)) {
                            h[l].t = tt; h[l].u = u; h[l].v = v; h[l].prim = prim; h[l].slot = slot;
                        }
                    }
                }
            }
        }
        if (any && tested) {
            for (int l = 0; l < 64; ++l)
                if (h[l].prim >= 0) alive &= ~(1ull << l);
            if (!alive) return;
        }
        if (!more) return;
        cur = next;
        m = next_m;
    }
}

/* The device's wave tiling: a 256-thread block covers 16x16 pixels, wave w of the block the 8x8
 * sub-tile ((w&1)*8, (w>>1)*8), lane l the pixel (l&7, l>>3) of it; every 8x8-aligned tile is one
 * wave. */
static inline void tile_lane(int64_t tile, uint32_t W, int64_t* col, int64_t* row, int l) {
    int64_t tiles_x8 = ((int64_t)W + 7) / 8;
    *col = (tile % tiles_x8) * 8 + (l & 7);
    *row = (tile / tiles_x8) * 8 + (l >> 3);
}

/* Camera closest hit with the device's wave-coherent traversal (rt_trace_camera). Per-pixel
 * outputs (may be NULL); wave_nodes / wave_leaves / wave_tris (may be NULL): 4-wide node records,
 * leaves and triangle records fetched per 8x8 tile (the device kernel's traffic and step count);
 * lane_counts (may be NULL): per pixel, the node records and triangles its own ray tested (the
 * nodes and leaves where its lane was in the wave's mask). */
TW_API void tw_trace_camera(const float* nodes4, const float* tris, int32_t root, const float* cam,
                            uint32_t W, uint32_t H, float* out_t, float* out_u, float* out_v,
                            int32_t* out_prim, int32_t* out_slot, int64_t* wave_nodes,
                            int64_t* wave_leaves, int64_t* wave_tris, int32_t* lane_counts,
                            int quad) {
    /* quad != 0: the quad-lane kernel's 16-ray packets over 4x4 tiles (ray q -> (q & 3, q >> 2)) */
    int64_t ts = quad ? 4 : 8, nr = quad ? 16 : 64;
    int64_t tiles_x = ((int64_t)W + ts - 1) / ts;
    int64_t ntiles = tiles_x * (((int64_t)H + ts - 1) / ts);
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t tile = 0; tile < ntiles; ++tile) {
        ray32 r[64];
        int want[64];
        int32_t skip[64];
        hit32 h[64];
        int64_t col[64], row[64];
        for (int l = 0; l < 64; ++l) {
            if (quad) {
                col[l] = (tile % tiles_x) * 4 + (l & 3);
                row[l] = (tile / tiles_x) * 4 + (l >> 2);
            } else {
                tile_lane(tile, W, &col[l], &row[l], l);
            }
            want[l] = l < nr && col[l] < W && row[l] < H;
            skip[l] = -1;
            float d[3];
            gen_ray(cam, W, H, want[l] ? (uint32_t)col[l] : 0, want[l] ? (uint32_t)row[l] : 0, d);
            ray32 rr = {cam[0], cam[1], cam[2], d[0], d[1], d[2], 0.0f, INFINITY};
            r[l] = rr;
        }
        wavecount wc = {0, 0, 0};
        packet32(nodes4, tris, root, r, want, skip, 0, h, &wc);
        if (wave_nodes) wave_nodes[tile] = wc.node_visits;
        if (wave_leaves) wave_leaves[tile] = wc.leaf_visits;
        if (wave_tris) wave_tris[tile] = wc.tri_fetches;
        for (int l = 0; l < 64; ++l) {
            if (!want[l]) continue;
            if (h[l].prim >= 0) refine64(tris, &r[l], &h[l]);
            int64_t i = row[l] * W + col[l];
            int hit = h[l].prim >= 0;
            if (out_t) out_t[i] = hit ? h[l].t : INFINITY;
            if (out_u) out_u[i] = hit ? h[l].u : 0.0f;
            if (out_v) out_v[i] = hit ? h[l].v : 0.0f;
            if (out_prim) out_prim[i] = h[l].prim;
            if (lane_counts) { lane_counts[2 * i] = h[l].node_visits; lane_counts[2 * i + 1] = h[l].tri_tests; }
            if (out_slot) out_slot[i] = h[l].slot;
        }
    }
}

/* Full FP32 render of the device path (rt_render): per 8x8 tile, wave-coherent primary closest
 * hit, FP64 hit refinement, Lambert, wave-coherent shadow any-hit of the lit hits, texture.
 * textures: concatenated RGBA8 images; tex_off/tex_w/tex_h per texture; mesh_tex per mesh.
 * params: light[3], ambient, background[3], shadow_tmin; flags bit0 shadows, bit1 bilinear. */
TW_API void tw_render(const float* nodes4, const float* tris, const float* shade, int32_t root,
                      const float* cam, uint32_t W, uint32_t H, const float* params, uint32_t flags,
                      const uint8_t* textures, const int64_t* tex_off, const int32_t* tex_w,
                      const int32_t* tex_h, const int32_t* mesh_tex, float* rgb, float* out_t,
                      int32_t* out_prim, int64_t* stats) {
    int64_t ntiles = (((int64_t)W + 7) / 8) * (((int64_t)H + 7) / 8);
    int64_t hits = 0, shadow_rays = 0, occl = 0;
    const float* L = params;
    #pragma omp parallel for schedule(dynamic, 16) reduction(+ : hits, shadow_rays, occl)
    for (int64_t tile = 0; tile < ntiles; ++tile) {
        ray32 r[64], sr[64];
        int want[64], wsh[64];
        int32_t skip[64], sskip[64];
        hit32 h[64], sh[64];
        int64_t col[64], row[64];
        float ndl[64], tu[64], tv[64];
        int32_t mesh[64];
        for (int l = 0; l < 64; ++l) {
            tile_lane(tile, W, &col[l], &row[l], l);
            want[l] = col[l] < W && row[l] < H;
            skip[l] = -1;
            float d[3];
            gen_ray(cam, W, H, want[l] ? (uint32_t)col[l] : 0, want[l] ? (uint32_t)row[l] : 0, d);
            ray32 rr = {cam[0], cam[1], cam[2], d[0], d[1], d[2], 0.0f, INFINITY};
            r[l] = rr;
        }
        packet32(nodes4, tris, root, r, want, skip, 0, h, NULL);
        for (int l = 0; l < 64; ++l) {
            int hit = want[l] && h[l].prim >= 0;
            ndl[l] = 0.0f; tu[l] = 0.0f; tv[l] = 0.0f; mesh[l] = 0;
            if (hit) {
                refine64(tris, &r[l], &h[l]);
                const float* s = shade + (int64_t)h[l].slot * 16;
                float w0 = 1.0f - h[l].u - h[l].v;
                float nx = fmaf(w0, s[0], fmaf(h[l].u, s[3], h[l].v * s[6]));
                float ny = fmaf(w0, s[1], fmaf(h[l].u, s[4], h[l].v * s[7]));
                float nz = fmaf(w0, s[2], fmaf(h[l].u, s[5], h[l].v * s[8]));
                float l2 = fmaf(nx, nx, fmaf(ny, ny, nz * nz));
                float il = l2 > 0.0f ? 1.0f / sqrtf(l2) : 0.0f;
                float nl = fmaf(nx * il, L[0], fmaf(ny * il, L[1], (nz * il) * L[2]));
                ndl[l] = fmaxf(nl, 0.0f);
                tu[l] = fmaf(w0, s[9], fmaf(h[l].u, s[11], h[l].v * s[13]));
                tv[l] = fmaf(w0, s[10], fmaf(h[l].u, s[12], h[l].v * s[14]));
                mesh[l] = (int32_t)fbits(s[15]);
            }
            wsh[l] = hit && (flags & 1u) && ndl[l] > 0.0f;
            sskip[l] = hit ? h[l].slot : -1;
            sr[l].ox = fmaf(h[l].t, r[l].dx, r[l].ox);
            sr[l].oy = fmaf(h[l].t, r[l].dy, r[l].oy);
            sr[l].oz = fmaf(h[l].t, r[l].dz, r[l].oz);
            sr[l].dx = L[0]; sr[l].dy = L[1]; sr[l].dz = L[2];
            sr[l].tmin = params[7]; sr[l].tmax = INFINITY;
        }
        packet32(nodes4, tris, root, sr, wsh, sskip, 1, sh, NULL);
        for (int l = 0; l < 64; ++l) {
            if (!want[l]) continue;
            int64_t i = row[l] * W + col[l];
            int hit = h[l].prim >= 0;
            int occ = wsh[l] && sh[l].prim >= 0;
            if (hit) hits++;
            if (wsh[l]) shadow_rays++;
            if (occ) occl++;
            if (out_t) out_t[i] = hit ? h[l].t : INFINITY;
            if (out_prim) out_prim[i] = hit ? h[l].prim : -1;
            if (!hit) {
                for (int c = 0; c < 3; ++c) rgb[i * 3 + c] = params[4 + c];
                continue;
            }
            float alb[3] = {1.0f, 1.0f, 1.0f};
            int32_t ti = mesh_tex ? mesh_tex[mesh[l]] : -1;
            if (ti >= 0)
                tex_sample(textures + tex_off[ti], tex_w[ti], tex_h[ti], tu[l], tv[l],
                           (flags & 2u) != 0, alb);
            float k = fmaf(occ ? 0.0f : 1.0f, ndl[l], params[3]);
            for (int c = 0; c < 3; ++c) rgb[i * 3 + c] = alb[c] * k;
        }
    }
    if (stats) { stats[0] = hits; stats[1] = shadow_rays; stats[2] = occl; }
}

/* ------------------------------------------------------------------------------------------ */
/* FP64 oracle (independent arithmetic)                                                         */
/* ------------------------------------------------------------------------------------------ */

/* FP64 Moller-Trumbore on FP32 data (converted exactly to double), with forward error bounds of
 * the FP32 evaluation of the same formula (mt32): err_* bound |fp32 result - exact result| using
 * the magnitudes of the intermediate products (unit roundoff 2^-24, safety factor 4). */
typedef struct {
    double t, u, v, det;
    double err_u, err_v, err_w, err_t, err_det;
} mt64_t;

static inline double norm3(const double* a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

static inline int mt64(const double* v0, const double* v1, const double* v2, const double* o,
                       const double* d, mt64_t* r) {
    /* exact edges: differences of FP32 values are exact in FP64 */
    double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2],
                   d[0] * e2[1] - d[1] * e2[0]};
    double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (det == 0.0) return 0;
    double s[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2],
                   s[0] * e1[1] - s[1] * e1[0]};
    double un = s[0] * p[0] + s[1] * p[1] + s[2] * p[2];
    double vn = d[0] * q[0] + d[1] * q[1] + d[2] * q[2];
    double tn = e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2];
    r->det = det;
    r->u = un / det;
    r->v = vn / det;
    r->t = tn / det;
    const double eps = 0x1p-24, K = 4.0;
    double nd = norm3(d), ne1 = norm3(e1), ne2 = norm3(e2), ns = norm3(s);
    double ad = fabs(det);
    r->err_det = K * eps * 5.0 * ne1 * nd * ne2;
    double e_un = K * eps * 6.0 * ns * nd * ne2;
    double e_vn = K * eps * 6.0 * nd * ns * ne1;
    double e_tn = K * eps * 6.0 * ne2 * ns * ne1;
    if (ad <= r->err_det) {
        r->err_u = r->err_v = r->err_w = r->err_t = INFINITY;
        return 1;
    }
    double den = ad - r->err_det;
    r->err_u = (e_un + fabs(r->u) * r->err_det) / den + K * eps * fabs(r->u);
    r->err_v = (e_vn + fabs(r->v) * r->err_det) / den + K * eps * fabs(r->v);
    r->err_w = r->err_u + r->err_v + K * eps;
    r->err_t = (e_tn + fabs(r->t) * r->err_det) / den + K * eps * fabs(r->t);
    return 1;
}

/* tri_pos: n_tris*9 floats in global primitive order (vertex k of prim p at p*9+k*3).
 * For each ray: out_t (+inf on miss), out_prim (-1 on miss), out_u/out_v, and out_edge, a bit set
 * of the reasons an FP32 evaluation may legitimately pick a different triangle:
 *   1  edge:      a barycentric of the winner is below max(edge_eps, its FP32 error bound)
 *   2  depth tie: another triangle is hit within max(tie_eps * t, the FP32 t error bounds of
 *                 both) of the winner's t (FP32 cannot order hits closer than its t accuracy)
 *   4  near miss: a triangle missed by less than max(edge_eps, its FP32 barycentric error bound)
 *                 lies at a t the FP32 evaluation could prefer (anywhere if nothing is hit) */
static void closest64_ray(const float* tri_pos, int64_t n_tris, const double* o, const double* d,
                          double tmin, double edge_eps, double tie_eps, const int64_t* cand,
                          int64_t n_cand, double* bt, int32_t* bp, double* bu, double* bv,
                          uint8_t* edge) {
    double best = INFINITY, best_err = 0.0;
    int32_t prim = -1;
    mt64_t w = {0};
    int64_t cnt = cand ? n_cand : n_tris;
    /* pass 1: the FP64 winner (ties in t: smaller primitive id) */
    for (int64_t k = 0; k < cnt; ++k) {
        int64_t p = cand ? cand[k] : k;
        const float* f = tri_pos + p * 9;
        double v0[3] = {f[0], f[1], f[2]}, v1[3] = {f[3], f[4], f[5]}, v2[3] = {f[6], f[7], f[8]};
        mt64_t r;
        if (!mt64(v0, v1, v2, o, d, &r)) continue;
        if (!(r.t > tmin)) continue;
        if (r.u < 0.0 || r.v < 0.0 || r.u + r.v > 1.0) continue;
        if (r.t < best || (r.t == best && (int32_t)p < prim)) {
            best = r.t; prim = (int32_t)p; w = r;
        }
    }
    *bt = best; *bp = prim; *bu = w.u; *bv = w.v;
    uint8_t e = 0;
    if (prim >= 0) {
        best_err = w.err_t;
        if (w.u < fmax(edge_eps, w.err_u) || w.v < fmax(edge_eps, w.err_v) ||
            1.0 - w.u - w.v < fmax(edge_eps, w.err_w))
            e |= 1;
    }
    /* pass 2: competitors the FP32 evaluation cannot separate from the winner */
    for (int64_t k = 0; k < cnt; ++k) {
        int64_t p = cand ? cand[k] : k;
        if ((int32_t)p == prim) continue;
        const float* f = tri_pos + p * 9;
        double v0[3] = {f[0], f[1], f[2]}, v1[3] = {f[3], f[4], f[5]}, v2[3] = {f[6], f[7], f[8]};
        mt64_t r;
        if (!mt64(v0, v1, v2, o, d, &r)) continue;
        double reach = prim >= 0 ? best + fmax(tie_eps * best, r.err_t + best_err) : INFINITY;
        if (!(r.t + r.err_t > tmin) || !(r.t <= reach)) continue;
        int inside = r.u >= 0.0 && r.v >= 0.0 && r.u + r.v <= 1.0;
        if (inside) {
            e |= 2;
        } else if (r.u > -fmax(edge_eps, r.err_u) && r.v > -fmax(edge_eps, r.err_v) &&
                   1.0 - r.u - r.v > -fmax(edge_eps, r.err_w)) {
            e |= 4;
        }
    }
    *edge = e;
}

/* FP64 box of a traversal ref: internal node box, or the union of a leaf's slot boxes. */
static void ref_box64(const double* nb, const double* lb, int32_t ref, double* b) {
    if (ref >= 0) {
        for (int c = 0; c < 6; ++c) b[c] = nb[(int64_t)ref * 6 + c];
        return;
    }
    uint32_t first = leaf_first(ref), count = leaf_count(ref);
    for (int c = 0; c < 3; ++c) { b[c] = INFINITY; b[3 + c] = -INFINITY; }
    for (uint32_t k = 0; k < count; ++k)
        for (int c = 0; c < 3; ++c) {
            b[c] = fmin(b[c], lb[(int64_t)(first + k) * 6 + c]);
            b[3 + c] = fmax(b[3 + c], lb[(int64_t)(first + k) * 6 + 3 + c]);
        }
}

/* FP64 slab test of a box already inflated by the margin. */
static int box_hit64(const double* b, const double* o, const double* d, double tmin, double margin) {
    double tn = -INFINITY, tf = INFINITY;
    for (int c = 0; c < 3; ++c) {
        if (d[c] == 0.0) {
            if (o[c] < b[c] || o[c] > b[3 + c]) return 0;
            continue;
        }
        double t0 = (b[c] - o[c]) / d[c], t1 = (b[3 + c] - o[c]) / d[c];
        if (t0 > t1) { double tt = t0; t0 = t1; t1 = tt; }
        tn = fmax(tn, t0); tf = fmin(tf, t1);
    }
    return !(tn > tf * (1.0 + 1e-12) + 1e-300 || tf < tmin - margin);
}

/* FP64 closest hit over every triangle whose box, inflated by `margin`, the ray crosses at
 * t >= tmin - margin (box_hit64). This is the candidate set the hierarchy version below reaches
 * through the tree: a hit lies inside its triangle's box, so no winner or depth tie is ever left
 * out, and a traversal only tests triangles of boxes its ray enters, so near misses are counted
 * among those triangles only (a nearly parallel triangle far off the ray has unbounded FP32 error
 * bounds but is never tested). */
TW_API void tw_closest64_brute(const float* tri_pos, int64_t n_tris, int64_t n, const float* ox,
                               const float* oy, const float* oz, const float* dx,
                               const float* dy, const float* dz, double tmin, double edge_eps,
                               double tie_eps, double margin, double* out_t, int32_t* out_prim,
                               double* out_u, double* out_v, uint8_t* out_edge) {
    double* tb = (double*)malloc(sizeof(double) * 6 * (n_tris > 0 ? n_tris : 1));
    for (int64_t p = 0; p < n_tris; ++p) {
        const float* f = tri_pos + p * 9;
        for (int c = 0; c < 3; ++c) {
            tb[p * 6 + c] = fmin(fmin((double)f[c], (double)f[3 + c]), (double)f[6 + c]) - margin;
            tb[p * 6 + 3 + c] = fmax(fmax((double)f[c], (double)f[3 + c]), (double)f[6 + c]) + margin;
        }
    }
    #pragma omp parallel
    {
        int64_t* cand = (int64_t*)malloc(sizeof(int64_t) * (n_tris > 0 ? n_tris : 1));
        #pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < n; ++i) {
            double o[3] = {ox[i], oy[i], oz[i]}, d[3] = {dx[i], dy[i], dz[i]};
            int64_t nc = 0;
            for (int64_t p = 0; p < n_tris; ++p)
                if (box_hit64(tb + p * 6, o, d, tmin, margin)) cand[nc++] = p;
            closest64_ray(tri_pos, n_tris, o, d, tmin, edge_eps, tie_eps, cand, nc, &out_t[i],
                          &out_prim[i], &out_u[i], &out_v[i], &out_edge[i]);
        }
        free(cand);
    }
    free(tb);
}

/* FP64 closest hit culled by an FP64 BVH over the same leaves: boxes are recomputed in FP64 from
 * the vertices and inflated by `margin` (absolute), so culling is conservative by construction.
 * nodes/order: the twin's hierarchy (only child refs are used). */
TW_API void tw_closest64_bvh(const float* tri_pos, int64_t n_tris, const float* nodes,
                             const uint32_t* order, int32_t root, int64_t n, const float* ox,
                             const float* oy, const float* oz, const float* dx, const float* dy,
                             const float* dz, double tmin, double edge_eps, double tie_eps,
                             double margin, double* out_t, int32_t* out_prim, double* out_u, double* out_v,
                             uint8_t* out_edge) {
    if (root == 0x7FFFFFFF || n_tris == 0) {
        for (int64_t i = 0; i < n; ++i) {
            out_t[i] = INFINITY; out_prim[i] = -1; out_u[i] = out_v[i] = 0; out_edge[i] = 0;
        }
        return;
    }
    int64_t nn = n_tris > 1 ? n_tris - 1 : 0;
    double* nb = (double*)malloc(sizeof(double) * 6 * (nn + 1));
    double* lb = (double*)malloc(sizeof(double) * 6 * n_tris);
    /* leaf boxes by slot */
    for (int64_t s = 0; s < n_tris; ++s) {
        const float* f = tri_pos + (int64_t)order[s] * 9;
        for (int c = 0; c < 3; ++c) {
            double lo = fmin(fmin((double)f[c], (double)f[3 + c]), (double)f[6 + c]);
            double hi = fmax(fmax((double)f[c], (double)f[3 + c]), (double)f[6 + c]);
            lb[s * 6 + c] = lo - margin;
            lb[s * 6 + 3 + c] = hi + margin;
        }
    }
    /* internal boxes by post-order */
    if (nn > 0) {
        int32_t* st = (int32_t*)malloc(sizeof(int32_t) * 2 * (n_tris + 1));
        uint8_t* state = (uint8_t*)calloc((size_t)nn, 1);
        int64_t sp = 0;
        st[sp++] = 0;
        while (sp) {
            int32_t x = st[sp - 1];
            int32_t l = (int32_t)fbits(nodes[x * 16 + 12]), r = (int32_t)fbits(nodes[x * 16 + 13]);
            if (!state[x]) {
                state[x] = 1;
                if (r >= 0) st[sp++] = r;
                if (l >= 0) st[sp++] = l;
                continue;
            }
            --sp;
            double bl[6], br[6];
            ref_box64(nb, lb, l, bl);
            ref_box64(nb, lb, r, br);
            for (int c = 0; c < 3; ++c) {
                nb[x * 6 + c] = fmin(bl[c], br[c]);
                nb[x * 6 + 3 + c] = fmax(bl[3 + c], br[3 + c]);
            }
        }
        free(st); free(state);
    }
    #pragma omp parallel
    {
        int64_t cap = 4096;
        int64_t* cand = (int64_t*)malloc(sizeof(int64_t) * cap);
        int32_t* stk = (int32_t*)malloc(sizeof(int32_t) * (n_tris + 64));
        #pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < n; ++i) {
            double o[3] = {ox[i], oy[i], oz[i]}, d[3] = {dx[i], dy[i], dz[i]};
            int64_t nc = 0, sp = 0;
            stk[sp++] = root;
            while (sp) {
                int32_t x = stk[--sp];
                if (x >= 0) {
                    if (!box_hit64(nb + (int64_t)x * 6, o, d, tmin, margin)) continue;
                    stk[sp++] = (int32_t)fbits(nodes[(int64_t)x * 16 + 12]);
                    stk[sp++] = (int32_t)fbits(nodes[(int64_t)x * 16 + 13]);
                } else {
                    uint32_t first = leaf_first(x), count = leaf_count(x);
                    for (uint32_t k = 0; k < count; ++k) {
                        if (!box_hit64(lb + (int64_t)(first + k) * 6, o, d, tmin, margin)) continue;
                        if (nc == cap) { cap *= 2; cand = (int64_t*)realloc(cand, sizeof(int64_t) * cap); }
                        cand[nc++] = order[first + k];
                    }
                }
            }
            closest64_ray(tri_pos, n_tris, o, d, tmin, edge_eps, tie_eps, cand, nc, &out_t[i],
                          &out_prim[i], &out_u[i], &out_v[i], &out_edge[i]);
        }
        free(cand); free(stk);
    }
    free(nb); free(lb);
}

TW_API int tw_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

TW_API void tw_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------------------------ */
/* Diffuse paths (rt_trace_paths), operation for operation                                   */
/* ------------------------------------------------------------------------------------------ */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds of two 32x32->64 multiplies,
 * key bumped by the Weyl constants between rounds. */
static void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int i = 0; i < 10; ++i) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1, n3 = (uint32_t)p0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

TW_API void tw_philox4x32_10(int64_t n, const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    for (int64_t i = 0; i < n; ++i) {
        uint32_t c[4] = {ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3]};
        philox4x32_10(c, key[0], key[1]);
        for (int k = 0; k < 4; ++k) out[4 * i + k] = c[k];
    }
}

static inline float unit_f(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

/* cos / sin of 2 pi x, x in [0,1): quadrant + Taylor polynomials on [0, pi/2) (Horner, fmaf) */
static void cos_sin_2pi(float x, float* c, float* s) {
    float x4 = x * 4.0f;
    int q = (int)x4;
    float a = (x4 - (float)q) * 1.57079632679489662f;
    float a2 = a * a;
    float sp = fmaf(a2, -2.50521083854417188e-8f, 2.75573192239858907e-6f);
    sp = fmaf(a2, sp, -1.98412698412698413e-4f);
    sp = fmaf(a2, sp, 8.33333333333333333e-3f);
    sp = fmaf(a2, sp, -1.66666666666666667e-1f);
    sp = fmaf(a2, sp, 1.0f);
    float sa = a * sp;
    float cp = fmaf(a2, 2.08767569878680990e-9f, -2.75573192239858907e-7f);
    cp = fmaf(a2, cp, 2.48015873015873016e-5f);
    cp = fmaf(a2, cp, -1.38888888888888889e-3f);
    cp = fmaf(a2, cp, 4.16666666666666667e-2f);
    cp = fmaf(a2, cp, -0.5f);
    float ca = fmaf(a2, cp, 1.0f);
    *c = q == 0 ? ca : q == 1 ? -sa : q == 2 ? -ca : sa;
    *s = q == 0 ? sa : q == 1 ? ca : q == 2 ? -sa : -ca;
}

TW_API void tw_cos_sin_2pi(int64_t n, const float* x, float* c, float* s) {
    for (int64_t i = 0; i < n; ++i) cos_sin_2pi(x[i], &c[i], &s[i]);
}

/* cosine-weighted direction about unit n: Malley's method in the Duff et al. 2017 basis */
static void cosine_dir(const float n[3], float u0, float u1, float d[3]) {
    float rad = sqrtf(u0), cs, sn;
    cos_sin_2pi(u1, &cs, &sn);
    float lx = rad * cs, ly = rad * sn, lz = sqrtf(fmaxf(0.0f, 1.0f - u0));
    float sg = copysignf(1.0f, n[2]);
    float a = -1.0f / (sg + n[2]);
    float b = n[0] * n[1] * a;
    float tx = fmaf(sg * n[0] * n[0], a, 1.0f), ty = sg * b, tz = -sg * n[0];
    float bx = b, by = fmaf(n[1] * n[1], a, sg), bz = -n[1];
    float x = fmaf(lx, tx, fmaf(ly, bx, lz * n[0]));
    float y = fmaf(lx, ty, fmaf(ly, by, lz * n[1]));
    float z = fmaf(lx, tz, fmaf(ly, bz, lz * n[2]));
    float il = 1.0f / sqrtf(fmaf(x, x, fmaf(y, y, z * z)));
    d[0] = x * il; d[1] = y * il; d[2] = z * il;
}

TW_API void tw_cosine_dir(int64_t n, const float* nrm, const float* u0, const float* u1, float* d) {
    for (int64_t i = 0; i < n; ++i) cosine_dir(nrm + 3 * i, u0[i], u1[i], d + 3 * i);
}

/* unit geometric normal e1 x e2 of a triangle record, facing against direction d */
static void facing_normal(const float* tr, const float d[3], float n[3]) {
    float x = fmaf(tr[5], tr[10], -(tr[6] * tr[9]));
    float y = fmaf(tr[6], tr[8], -(tr[4] * tr[10]));
    float z = fmaf(tr[4], tr[9], -(tr[5] * tr[8]));
    float il = 1.0f / sqrtf(fmaf(x, x, fmaf(y, y, z * z)));
    x *= il; y *= il; z *= il;
    int flip = fmaf(x, d[0], fmaf(y, d[1], z * d[2])) > 0.0f;
    n[0] = flip ? -x : x; n[1] = flip ? -y : y; n[2] = flip ? -z : z;
}

/* rt_trace_paths: per pixel and sample, the jittered camera ray and up to `bounces` diffuse
 * bounces (closest hits by the per-ray traversal, FP64-refined); samples[pixel*spp+s] = sky *
 * albedo^k for a path escaping after k bounces, else 0; radiance = mean over samples in order;
 * counts[k] = rays traced at depth k (k = 0..bounces). */
TW_API void tw_paths(const float* nodes, const float* tris, int32_t root, const float* cam,
                     uint32_t W, uint32_t H, uint32_t spp, uint32_t bounces, uint32_t key0,
                     uint32_t key1, float albedo, float sky, float tmin, float* samples,
                     float* radiance, int64_t* counts) {
    int64_t npx = (int64_t)W * H;
    int64_t cnt[17] = {0};
    #pragma omp parallel for schedule(dynamic, 64) reduction(+:cnt[0:17])
    for (int64_t p = 0; p < npx; ++p) {
        uint32_t col = (uint32_t)(p % W), row = (uint32_t)(p / W);
        float sum = 0.0f;
        for (uint32_t s = 0; s < spp; ++s) {
            uint32_t c[4] = {(uint32_t)p, s, 0u, 0u};
            philox4x32_10(c, key0, key1);
            float fx = (float)col + unit_f(c[0]), fy = (float)row + unit_f(c[1]);
            float nx = fx / (float)W * 2.0f - 1.0f, ny = 1.0f - fy / (float)H * 2.0f;
            float x = fmaf(nx, cam[3], fmaf(ny, cam[6], cam[9]));
            float y = fmaf(nx, cam[4], fmaf(ny, cam[7], cam[10]));
            float z = fmaf(nx, cam[5], fmaf(ny, cam[8], cam[11]));
            float il = 1.0f / sqrtf(fmaf(x, x, fmaf(y, y, z * z)));
            ray32 r = {cam[0], cam[1], cam[2], x * il, y * il, z * il, 0.0f, INFINITY};
            hit32 h;
            traverse32(nodes, tris, root, &r, 0, -1, &h);
            cnt[0]++;
            float val = 0.0f, thr = 1.0f;
            if (h.prim < 0) {
                val = sky;
            } else {
                refine64(tris, &r, &h);
                for (uint32_t k = 1; k <= bounces; ++k) {
                    const float* tr = tris + (int64_t)h.slot * 12;
                    float d[3] = {r.dx, r.dy, r.dz}, n[3], nd[3];
                    facing_normal(tr, d, n);
                    uint32_t c2[4] = {(uint32_t)p, s, k, 0u};
                    philox4x32_10(c2, key0, key1);
                    cosine_dir(n, unit_f(c2[0]), unit_f(c2[1]), nd);
                    float ox = fmaf(h.t, r.dx, r.ox), oy = fmaf(h.t, r.dy, r.oy), oz = fmaf(h.t, r.dz, r.oz);
                    thr = thr * albedo;
                    int32_t skip = h.slot;
                    ray32 rb = {ox, oy, oz, nd[0], nd[1], nd[2], tmin, INFINITY};
                    r = rb;
                    traverse32(nodes, tris, root, &r, 0, skip, &h);
                    cnt[k]++;
                    if (h.prim < 0) { val = thr * sky; break; }
                    refine64(tris, &r, &h);
                }
            }
            if (samples) samples[p * spp + s] = val;
            sum += val;
        }
        radiance[p] = sum / (float)spp;
    }
    for (uint32_t k = 0; k <= bounces && k < 17; ++k) counts[k] = cnt[k];
}

/* ---- per-vertex normals: Mesh.compute_normals (rendering/_modeling.py:48-49, a stub) ----
 * n(v) = sum of cross(P1 - P0, P2 - P0) over the (triangle t, corner k) pairs with idx[t][k] == v,
 * added in the order k = 0, 1, 2 and ascending t within k, starting from +0; N(v) = n / d with
 * d = sqrt((x*x + y*y) + z*z) raised to 1e-300 (a NaN d stays NaN); FP64 from the FP32 positions,
 * rounded once to FP32. Out-of-range indices are clamped to n_vertices - 1; idx == NULL means the
 * vertices 3t, 3t+1, 3t+2. pos: n_vertices x 3 floats; out: n_vertices x 3 floats. */
TW_API void tw_vertex_normals(const float* pos, int64_t n_vertices, const int32_t* idx, int64_t n_tris,
                              float* out) {
    double* acc = (double*)calloc((size_t)(3 * (n_vertices ? n_vertices : 1)), sizeof(double));
    for (int k = 0; k < 3; ++k) {
        for (int64_t t = 0; t < n_tris; ++t) {
            int64_t c[3];
            for (int j = 0; j < 3; ++j) {
                if (idx) {
                    const uint32_t u = (uint32_t)idx[3 * t + j];
                    c[j] = u < (uint64_t)n_vertices ? (int64_t)u : n_vertices - 1;
                } else {
                    c[j] = 3 * t + j;
                }
            }
            const float* a = pos + 3 * c[0];
            const float* b = pos + 3 * c[1];
            const float* d = pos + 3 * c[2];
            const double e1x = (double)b[0] - (double)a[0], e1y = (double)b[1] - (double)a[1],
                         e1z = (double)b[2] - (double)a[2];
            const double e2x = (double)d[0] - (double)a[0], e2y = (double)d[1] - (double)a[1],
                         e2z = (double)d[2] - (double)a[2];
            double* s = acc + 3 * c[k];
            s[0] = s[0] + (e1y * e2z - e1z * e2y);
            s[1] = s[1] + (e1z * e2x - e1x * e2z);
            s[2] = s[2] + (e1x * e2y - e1y * e2x);
        }
    }
    for (int64_t v = 0; v < n_vertices; ++v) {
        const double x = acc[3 * v], y = acc[3 * v + 1], z = acc[3 * v + 2];
        double d = sqrt((x * x + y * y) + z * z);
        d = d < 1e-300 ? 1e-300 : d;
        out[3 * v] = (float)(x / d);
        out[3 * v + 1] = (float)(y / d);
        out[3 * v + 2] = (float)(z / d);
    }
    free(acc);
}

/* ---- raster cross-check of primary visibility ----
 * A second statement of which triangle each pixel centre sees, decided the way the reference's
 * raster pipeline decides it rather than by rays: project the vertices, clip at the near plane
 * (rendering/_raster.py:172-214), test the pixel centre (col + 0.5, row + 0.5) against screen-space
 * edge functions (:282-311) and keep the nearest depth (:80-112). FP64 throughout, without the
 * reference's fragment-capacity and 64x64-pixel limits. Camera: the device's 12 floats (eye,
 * right, up, forward); a point P maps to (a, b, s) = M^-1 (P - eye) with M = [right up forward],
 * ndc = (a/s, b/s), X = (ndc.x + 1) W/2, Y = (1 - ndc.y) H/2: the inverse of the device's ray
 * generation, so the depth order along a pixel's ray is the order of s, and 1/s is affine in
 * (X, Y) over a triangle. Per pixel: prim = the triangle of smallest s whose closed screen
 * triangle contains the centre (ties to the smaller prim), -1 if none. Depths are compared as ranges: a triangle's 1/s over the disc of radius
 * edge_eps around the centre (its screen gradient times edge_eps), widened by depth_eps relative.
 * flags: 1 (edge) when a triangle whose edge passes within edge_eps pixels of the centre may lie at
 * or in front of the winner; 2 (depth tie) when another covering triangle may. Those are the pixels
 * on which a ray moved by less than edge_eps pixels could see another triangle. */
typedef struct {
    double X[3], Y[3], iz[3];
    double x0, x1, y0, y1;
    int32_t prim;
    int internal;  /* bit j: edge j (vertex j -> j+1) is a diagonal of a clipped polygon's fan */
    double giz;    /* |grad (1/s)| in screen space, per pixel */
} rtri;

static int clip_near(double v[][3], int n, double sn, double out[][3]) {
    int m = 0;
    for (int i = 0; i < n; ++i) {
        const double* p = v[i];
        const double* q = v[(i + 1) % n];
        const int pin = p[2] >= sn, qin = q[2] >= sn;
        if (pin) { out[m][0] = p[0]; out[m][1] = p[1]; out[m][2] = p[2]; ++m; }
        if (pin != qin) {
            const double f = (sn - p[2]) / (q[2] - p[2]);
            out[m][0] = p[0] + f * (q[0] - p[0]);
            out[m][1] = p[1] + f * (q[1] - p[1]);
            out[m][2] = sn;
            ++m;
        }
    }
    return m;
}

TW_API void tw_raster_ids(const float* tri_pos, int64_t n_tris, const float* cam, uint32_t W, uint32_t H,
                          double edge_eps, double depth_eps, int32_t* prim_out, uint8_t* flags_out) {
    /* M^-1 by the adjugate */
    double M[3][3];
    for (int r = 0; r < 3; ++r) {
        M[r][0] = cam[3 + r];
        M[r][1] = cam[6 + r];
        M[r][2] = cam[9 + r];
    }
    const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                       M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                       M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    double Mi[3][3];
    Mi[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
    Mi[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det;
    Mi[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
    Mi[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) / det;
    Mi[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
    Mi[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
    Mi[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) / det;
    Mi[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / det;
    Mi[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
    const double sn = 1e-9;
    rtri* rt = (rtri*)malloc(sizeof(rtri) * (size_t)(2 * (n_tris ? n_tris : 1)));
    int64_t nr = 0;
    for (int64_t t = 0; t < n_tris; ++t) {
        double v[3][3], c[4][3];
        for (int k = 0; k < 3; ++k) {
            const double d0 = (double)tri_pos[9 * t + 3 * k] - cam[0];
            const double d1 = (double)tri_pos[9 * t + 3 * k + 1] - cam[1];
            const double d2 = (double)tri_pos[9 * t + 3 * k + 2] - cam[2];
            for (int r = 0; r < 3; ++r) v[k][r] = Mi[r][0] * d0 + Mi[r][1] * d1 + Mi[r][2] * d2;
        }
        const int m = clip_near(v, 3, sn, c);
        for (int f = 1; f + 1 < m; ++f) {  /* fan of the clipped polygon */
            const int id[3] = {0, f, f + 1};
            rtri* q = &rt[nr];
            for (int k = 0; k < 3; ++k) {
                const double* p = c[id[k]];
                q->X[k] = (p[0] / p[2] + 1.0) * 0.5 * W;
                q->Y[k] = (1.0 - p[1] / p[2]) * 0.5 * H;
                q->iz[k] = 1.0 / p[2];
            }
            q->x0 = fmin(fmin(q->X[0], q->X[1]), q->X[2]);
            q->x1 = fmax(fmax(q->X[0], q->X[1]), q->X[2]);
            q->y0 = fmin(fmin(q->Y[0], q->Y[1]), q->Y[2]);
            q->y1 = fmax(fmax(q->Y[0], q->Y[1]), q->Y[2]);
            q->prim = (int32_t)t;
            q->internal = (f > 1 ? 1 : 0) | (f + 2 < m ? 4 : 0);
            {
                const double ar = (q->X[1] - q->X[0]) * (q->Y[2] - q->Y[0]) - (q->Y[1] - q->Y[0]) * (q->X[2] - q->X[0]);
                const double gx = ((q->iz[1] - q->iz[0]) * (q->Y[2] - q->Y[0]) - (q->iz[2] - q->iz[0]) * (q->Y[1] - q->Y[0])) / ar;
                const double gy = ((q->X[1] - q->X[0]) * (q->iz[2] - q->iz[0]) - (q->X[2] - q->X[0]) * (q->iz[1] - q->iz[0])) / ar;
                q->giz = hypot(gx, gy);
            }
            ++nr;
        }
    }
    const int64_t band = 8, nbands = ((int64_t)H + band - 1) / band;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < nbands; ++b) {
        const int64_t r0 = b * band, r1 = r0 + band < (int64_t)H ? r0 + band : (int64_t)H;
        const int64_t np = (r1 - r0) * W;
        /* per pixel: winner's iz, its lowest and highest iz within edge_eps, the highest iz within
           edge_eps of any other covering triangle, of any ambiguous one */
        double* best = (double*)malloc(sizeof(double) * 5 * (size_t)np);
        int32_t* bp = (int32_t*)malloc(sizeof(int32_t) * (size_t)np);
        for (int64_t i = 0; i < np; ++i) {
            for (int j = 0; j < 5; ++j) best[5 * i + j] = -INFINITY;
            bp[i] = -1;
        }
        for (int64_t k = 0; k < nr; ++k) {
            const rtri* q = &rt[k];
            if (q->y1 + edge_eps < r0 + 0.5 || q->y0 - edge_eps > r1 - 0.5) continue;
            const double area = (q->X[1] - q->X[0]) * (q->Y[2] - q->Y[0]) - (q->Y[1] - q->Y[0]) * (q->X[2] - q->X[0]);
            if (area == 0.0 || !(fabs(area) < INFINITY)) continue;
            const double sg = area > 0 ? 1.0 : -1.0;
            double len[3];
            for (int e = 0; e < 3; ++e) {
                const int e2 = (e + 1) % 3;
                len[e] = hypot(q->X[e2] - q->X[e], q->Y[e2] - q->Y[e]);
            }
            int64_t ca = (int64_t)ceil(q->x0 - edge_eps - 0.5), cb = (int64_t)floor(q->x1 + edge_eps - 0.5);
            int64_t ra = (int64_t)ceil(q->y0 - edge_eps - 0.5), rb = (int64_t)floor(q->y1 + edge_eps - 0.5);
            if (ca < 0) ca = 0;
            if (cb > (int64_t)W - 1) cb = (int64_t)W - 1;
            if (ra < r0) ra = r0;
            if (rb > r1 - 1) rb = r1 - 1;
            for (int64_t row = ra; row <= rb; ++row) {
                const double py = row + 0.5;
                for (int64_t col = ca; col <= cb; ++col) {
                    const double px = col + 0.5;
                    double e[3], dmin = INFINITY;
                    int beyond_diagonal = 0;
                    for (int j = 0; j < 3; ++j) {
                        const int j2 = (j + 1) % 3;
                        e[j] = (q->X[j2] - q->X[j]) * (py - q->Y[j]) - (q->Y[j2] - q->Y[j]) * (px - q->X[j]);
                        const double dist = sg * e[j] / len[j];
                        if ((q->internal >> j) & 1) {
                            /* a fan diagonal is no silhouette: the other half of the polygon covers
                               the centre (both halves are the same primitive) */
                            beyond_diagonal |= dist < -1e-9;
                        } else {
                            dmin = dist < dmin ? dist : dmin;
                        }
                    }
                    if (beyond_diagonal || dmin <= -edge_eps) continue;
                    /* barycentric of vertex k = edge function of the opposite edge / area */
                    const double iz = (e[1] * q->iz[0] + e[2] * q->iz[1] + e[0] * q->iz[2]) / area;
                    const double lo = iz - edge_eps * q->giz, hi = iz + edge_eps * q->giz;
                    const int64_t i = (row - r0) * W + col;
                    double* st = best + 5 * i;
                    if (dmin < edge_eps && hi > st[4]) st[4] = hi;  /* coverage ambiguous */
                    if (dmin < 0.0) continue;
                    if (bp[i] < 0 || iz > st[0] || (iz == st[0] && q->prim < bp[i])) {
                        if (bp[i] >= 0 && bp[i] != q->prim && st[2] > st[3]) st[3] = st[2];
                        st[0] = iz;
                        st[1] = lo;
                        st[2] = hi;
                        bp[i] = q->prim;
                    } else if (q->prim != bp[i] && hi > st[3]) {
                        st[3] = hi;
                    }
                }
            }
        }
        for (int64_t i = 0; i < np; ++i) {
            const double* st = best + 5 * i;
            const double floor_iz = st[1] * (1.0 - depth_eps);
            uint8_t fl = 0;
            if (st[4] > -INFINITY && (bp[i] < 0 || st[4] >= floor_iz)) fl |= 1;
            if (bp[i] >= 0 && st[3] > -INFINITY && st[3] >= floor_iz) fl |= 2;
            prim_out[r0 * W + i] = bp[i];
            flags_out[r0 * W + i] = fl;
        }
        free(best);
        free(bp);
    }
    free(rt);
}

/* ---- image export: float RGB -> BGRA8 with the OpenCL UNORM_INT8 write conversion
 * convert_uchar_sat_rte(f * 255) (rendering/_core.py:340; OpenCL 1.2 section 8.3.1.1), NaN -> 0,
 * alpha 255. */
static inline uint8_t unorm8(float f) {
    const float x = f * 255.0f;
    if (!(x > 0.0f)) return 0;
    if (x >= 255.0f) return 255;
    return (uint8_t)nearbyintf(x);  /* default rounding mode: ties to even */
}

TW_API void tw_rgb_to_bgra8(const float* rgb, int64_t n, uint8_t* out) {
    for (int64_t i = 0; i < n; ++i) {
        out[4 * i] = unorm8(rgb[3 * i + 2]);
        out[4 * i + 1] = unorm8(rgb[3 * i + 1]);
        out[4 * i + 2] = unorm8(rgb[3 * i]);
        out[4 * i + 3] = 255;
    }
}
