// rt_build.hip — on-device LBVH construction for rendering/_raycaster.py:26-33
// (Raycaster.__init__ -> _build_ads, a stub in the reference that only counts triangles).
//
// Pipeline (one stream; no host synchronisation except the PLOC depth check):
//   1. gather      per mesh: read MeshVertex P/N/C (rendering/_modeling.py:22-28) through the
//                  optional int32 index list, write unsorted tri/shade records, padded leaf box,
//                  AABB centroid.
//   2. bounds      two-stage min/max reduction of the centroids.
//   3. morton      30-bit Morton code of each centroid quantised to 1024^3 over those bounds.
//   4. sort        stable LSD radix sort of (code, primitive) in 4 passes of 8 bits (rt_sort.h):
//                  per-tile histogram -> per-digit row scan -> scatter, the scatter ranking keys
//                  with wave64 ballot/mbcnt multi-split so equal digits keep input order and
//                  writing each tile digit by digit from LDS.
//   5. hierarchy   Karras 2012 binary radix tree over the extended keys (code, slot).
//   6. emit        permute tri/shade records into slot order.
//   7. refit       nodes whose leaf range lies inside one 1024-leaf block: bottom-up in that
//                  block's workgroup (the second thread to reach a node unions its children's
//                  boxes and continues upward, LDS counters). The ~0.6% spanning blocks: each
//                  child box is the union of its slot range, from per-block prefix/suffix unions
//                  and a sparse table over blocks (min/max are exact, so this equals the
//                  bottom-up union), one thread per node with no waiting.
//   8. wide        4-wide records from the binary ones.
// PLOC (rt_tuning.builder) replaces 5-7 with locally ordered clustering over the same order.
// Every result is an integer or an exact FP32 min/max/sub, so the output is byte-identical to the
// CPU restatement under the same inputs.
#include <algorithm>
#include <cmath>
#include <utility>

#include "rt_common.h"
#include "rt_sort.h"

namespace rt {
namespace {

constexpr int kReduceBlocks = 512;

struct Workspace {
    float4* cent;      // n
    float4* leafbox;   // 2n
    float4* tri_u;     // 3n
    float4* shade_u;   // 4n
    uint32_t* keys0;   // n
    uint32_t* vals0;
    uint32_t* keys1;
    uint32_t* vals1;
    uint32_t* hist;    // kRadix * nblk
    uint32_t* totals;  // kRadix
    float* partial;    // kReduceBlocks * 8
    float* cbounds;    // 8
    int32_t* parent;   // (n-1) internal then n leaves
    int32_t* rangej;   // n-1: the other end j of internal node i's leaf range [min(i,j), max(i,j)]
    int32_t* split;    // n-1: split position g of internal node i (children ranges [lo,g], [g+1,hi])
    float4* nodebox;   // 2(n-1)
    int32_t* top_list; // n-1: internal nodes whose leaf range spans several refit blocks
    uint32_t* top_count;
    // PLOC builder
    int32_t* cid[2];   // n: cluster refs by position (internal id, or ~prim), ping-pong
    float4* cbox[2];   // 2n: cluster boxes by position, ping-pong
    int32_t* nn;       // n: nearest neighbour position
    int32_t* nleft;    // n-1: children and triangle count of each internal node
    int32_t* nright;
    int32_t* nsize;
    uint32_t* pos_of;  // n: Morton position of each primitive
    uint32_t* slot_of; // n: slot of each primitive
    uint32_t* fslot;   // n-1: first slot of each internal node inside a leaf
    uint32_t* scan;    // n: leaf-root sizes at their first Morton position, then their first slot
    uint32_t* part;    // 2 * chunks: per-chunk counts, then exclusive offsets
    int32_t* state;    // 8: PlocState per iteration parity
    uint32_t* depth;   // 1: internal ancestors of the deepest triangle of the PLOC tree
    float4* pre;       // 2n: union of the leaf boxes from the start of the slot's refit block to it
    float4* suf;       // 2n: union from the slot to the end of its refit block
    float4* sparse;    // 2 * nb * levels: unions of 2^k consecutive refit blocks
    uint32_t* slot_order;  // n: primitive of each slot
    uint64_t bytes;
};

inline uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

// Refit blocks: bottom-up refit inside each range of kRefitBlock leaves runs in one workgroup
// (k_refit_local, which also scans the block's leaf boxes into prefix/suffix unions); the nodes
// whose range spans several blocks ("top" nodes) are refit afterwards from those range unions
// (k_block_sparse, k_refit_top).
constexpr int kRefitBlock = 1024;

// levels of the sparse table over nb blocks: level k holds unions of 2^k consecutive blocks
inline uint32_t sparse_levels(uint64_t nb) {
    uint32_t l = 1;
    while ((1ull << l) <= nb) ++l;
    return l;
}

Workspace carve(char* base, uint64_t n) {
    Workspace w{};
    const uint64_t nblk = (n + kSortTile - 1) / kSortTile;
    const uint64_t ni = n > 1 ? n - 1 : 1;
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        char* p = base ? base + off : nullptr;
        off = align256(off + bytes);
        return p;
    };
    w.cent = (float4*)take(16 * n);
    w.leafbox = (float4*)take(32 * n);
    w.tri_u = (float4*)take(48 * n);
    w.shade_u = (float4*)take(64 * n);
    w.keys0 = (uint32_t*)take(4 * n);
    w.vals0 = (uint32_t*)take(4 * n);
    w.keys1 = (uint32_t*)take(4 * n);
    w.vals1 = (uint32_t*)take(4 * n);
    w.hist = (uint32_t*)take(4 * kRadix * (nblk ? nblk : 1));
    w.totals = (uint32_t*)take(4 * kRadix);
    w.partial = (float*)take(4 * 8 * kReduceBlocks);
    w.cbounds = (float*)take(4 * 8);
    w.parent = (int32_t*)take(4 * (ni + n));
    w.rangej = (int32_t*)take(4 * ni);
    w.split = (int32_t*)take(4 * ni);
    w.nodebox = (float4*)take(32 * ni);
    w.top_list = (int32_t*)take(4 * ni);
    w.top_count = (uint32_t*)take(4);
    const uint64_t nchunk = (n + 1023) / 1024 + 1;
    for (int k = 0; k < 2; ++k) {
        w.cid[k] = (int32_t*)take(4 * n);
        w.cbox[k] = (float4*)take(32 * n);
    }
    w.nn = (int32_t*)take(4 * n);
    w.nleft = (int32_t*)take(4 * ni);
    w.nright = (int32_t*)take(4 * ni);
    w.nsize = (int32_t*)take(4 * ni);
    w.pos_of = (uint32_t*)take(4 * n);
    w.slot_of = (uint32_t*)take(4 * n);
    w.fslot = (uint32_t*)take(4 * ni);
    w.scan = (uint32_t*)take(4 * n);
    w.part = (uint32_t*)take(8 * nchunk);
    w.state = (int32_t*)take(4 * 8);
    w.depth = (uint32_t*)take(4);
    w.pre = (float4*)take(32 * n);
    w.suf = (float4*)take(32 * n);
    {
        const uint64_t nb = (n + kRefitBlock - 1) / kRefitBlock;
        w.sparse = (float4*)take(32 * nb * sparse_levels(nb));
    }
    w.slot_order = (uint32_t*)take(4 * n);
    w.bytes = off;
    return w;
}

RT_DEV float canon(float x) { return x + 0.0f; }  // -0 -> +0

// Leaf boxes are padded by |coordinate| * 2^-14 + 2^-40 per side, wider than the world-space error
// of an FP32 Moller-Trumbore acceptance, so a box test never culls a triangle the FP32 test would
// accept: the closest hit is then independent of the traversal order (per-ray or wave-coherent).
RT_DEV float pad_of(float lo, float hi) {
    const float m = fmaxf(fabsf(lo), fabsf(hi));
    return m * 0x1p-14f + 0x1p-40f;
}

RT_DEV float3 load_f3(const char* base, uint64_t i, uint32_t stride, uint32_t off) {
    const float* p = (const float*)(base + i * stride + off);
    return make_float3(p[0], p[1], p[2]);
}

// Copies `per` float4s of each of the block's `rows` records from LDS (record r at
// stage[r * pitch]) to out[0 .. rows*per), one float4 per thread and instruction.
RT_DEV void store_staged(float4* __restrict__ out, const float4* stage, uint32_t rows, uint32_t per,
                         uint32_t pitch) {
    for (uint32_t q = threadIdx.x; q < rows * per; q += blockDim.x) {
        const uint32_t r = q / per;
        out[q] = stage[r * pitch + (q - r * per)];
    }
}

// One thread per triangle: positions (canonical zeros), the centroid, the padded leaf box, the
// intersection record (p0, e1, e2, primitive id) and the shading record (three normals, three
// uvs, mesh id). The block's records go out through LDS as consecutive float4s.
__global__ void __launch_bounds__(256) k_gather(rt_mesh_desc m, uint32_t mesh_id,
                                                 uint32_t prim_base, float4* __restrict__ cent,
                                                 float4* __restrict__ leafbox,
                                                 float4* __restrict__ tri_u,
                                                 float4* __restrict__ shade_u,
                                                 int32_t* __restrict__ mesh_offsets) {
    constexpr uint32_t kPitch = 5;  // float4s per staged record (4 used)
    __shared__ float4 stage[256 * kPitch];
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) mesh_offsets[mesh_id] = (int32_t)prim_base;
    const uint32_t t0 = blockIdx.x * blockDim.x;
    if (t0 >= m.n_tris) return;  // whole block (only the launch for an empty mesh)
    const uint32_t rows = m.n_tris - t0 < blockDim.x ? m.n_tris - t0 : blockDim.x;
    const uint64_t pb = (uint64_t)prim_base + t0;  // the block's first primitive
    const bool mine = t < m.n_tris;
    float4 box[2], tri[3], shd[4];
    if (mine) {
        const char* vb = (const char*)m.vertices;
        uint32_t i0 = 3 * t, i1 = 3 * t + 1, i2 = 3 * t + 2;
        if (m.indices) {
            i0 = (uint32_t)m.indices[3 * (uint64_t)t];
            i1 = (uint32_t)m.indices[3 * (uint64_t)t + 1];
            i2 = (uint32_t)m.indices[3 * (uint64_t)t + 2];
            // out-of-range indices are clamped so a malformed list cannot read past the buffer
            i0 = i0 < m.n_vertices ? i0 : m.n_vertices - 1;
            i1 = i1 < m.n_vertices ? i1 : m.n_vertices - 1;
            i2 = i2 < m.n_vertices ? i2 : m.n_vertices - 1;
        }
        float3 p0 = load_f3(vb, i0, m.vertex_stride, m.pos_offset);
        float3 p1 = load_f3(vb, i1, m.vertex_stride, m.pos_offset);
        float3 p2 = load_f3(vb, i2, m.vertex_stride, m.pos_offset);
        p0 = make_float3(canon(p0.x), canon(p0.y), canon(p0.z));
        p1 = make_float3(canon(p1.x), canon(p1.y), canon(p1.z));
        p2 = make_float3(canon(p2.x), canon(p2.y), canon(p2.z));
        const uint32_t p = prim_base + t;

        const float lox = fminf(fminf(p0.x, p1.x), p2.x), hix = fmaxf(fmaxf(p0.x, p1.x), p2.x);
        const float loy = fminf(fminf(p0.y, p1.y), p2.y), hiy = fmaxf(fmaxf(p0.y, p1.y), p2.y);
        const float loz = fminf(fminf(p0.z, p1.z), p2.z), hiz = fmaxf(fmaxf(p0.z, p1.z), p2.z);
        cent[p] = make_float4((lox + hix) * 0.5f, (loy + hiy) * 0.5f, (loz + hiz) * 0.5f, 0.0f);
        const float px = pad_of(lox, hix), py = pad_of(loy, hiy), pz = pad_of(loz, hiz);
        box[0] = make_float4(lox - px, loy - py, loz - pz, 0.0f);
        box[1] = make_float4(hix + px, hiy + py, hiz + pz, 0.0f);

        const float e1x = p1.x - p0.x, e1y = p1.y - p0.y, e1z = p1.z - p0.z;
        const float e2x = p2.x - p0.x, e2y = p2.y - p0.y, e2z = p2.z - p0.z;
        tri[0] = make_float4(p0.x, p0.y, p0.z, __int_as_float((int32_t)p));
        tri[1] = make_float4(e1x, e1y, e1z, 0.0f);
        tri[2] = make_float4(e2x, e2y, e2z, 0.0f);

        float3 n0, n1, n2;
        if (m.nrm_offset != RT_ABSENT) {
            n0 = load_f3(vb, i0, m.vertex_stride, m.nrm_offset);
            n1 = load_f3(vb, i1, m.vertex_stride, m.nrm_offset);
            n2 = load_f3(vb, i2, m.vertex_stride, m.nrm_offset);
        } else {
            const float gx = fmaf(e1y, e2z, -(e1z * e2y));
            const float gy = fmaf(e1z, e2x, -(e1x * e2z));
            const float gz = fmaf(e1x, e2y, -(e1y * e2x));
            const float l2 = fmaf(gx, gx, fmaf(gy, gy, gz * gz));
            const float il = l2 > 0.0f ? 1.0f / sqrtf(l2) : 0.0f;
            n0 = n1 = n2 = make_float3(gx * il, gy * il, gz * il);
        }
        float2 uv0 = make_float2(0.0f, 0.0f), uv1 = uv0, uv2 = uv0;
        if (m.uv_offset != RT_ABSENT) {
            const float* q0 = (const float*)(vb + (uint64_t)i0 * m.vertex_stride + m.uv_offset);
            const float* q1 = (const float*)(vb + (uint64_t)i1 * m.vertex_stride + m.uv_offset);
            const float* q2 = (const float*)(vb + (uint64_t)i2 * m.vertex_stride + m.uv_offset);
            uv0 = make_float2(q0[0], q0[1]);
            uv1 = make_float2(q1[0], q1[1]);
            uv2 = make_float2(q2[0], q2[1]);
        }
        shd[0] = make_float4(n0.x, n0.y, n0.z, n1.x);
        shd[1] = make_float4(n1.y, n1.z, n2.x, n2.y);
        shd[2] = make_float4(n2.z, uv0.x, uv0.y, uv1.x);
        shd[3] = make_float4(uv1.y, uv2.x, uv2.y, __int_as_float((int32_t)mesh_id));
    }
    float4* st = stage + threadIdx.x * kPitch;
    if (mine) { st[0] = shd[0]; st[1] = shd[1]; st[2] = shd[2]; st[3] = shd[3]; }
    __syncthreads();
    store_staged(shade_u + 4 * pb, stage, rows, 4, kPitch);
    __syncthreads();
    if (mine) { st[0] = tri[0]; st[1] = tri[1]; st[2] = tri[2]; }
    __syncthreads();
    store_staged(tri_u + 3 * pb, stage, rows, 3, kPitch);
    __syncthreads();
    if (mine) { st[0] = box[0]; st[1] = box[1]; }
    __syncthreads();
    store_staged(leafbox + 2 * pb, stage, rows, 2, kPitch);
}

__global__ void k_mesh_offsets_tail(int32_t* mesh_offsets, uint32_t n_meshes, uint32_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) mesh_offsets[n_meshes] = (int32_t)n;
}

// Stage 1: each block reduces a grid-stride share of the centroids to (min.xyz, max.xyz).
__global__ void __launch_bounds__(256) k_bounds_partial(const float4* __restrict__ cent,
                                                         uint32_t n, float* __restrict__ partial) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 c = cent[i];
        lo[0] = fminf(lo[0], c.x); lo[1] = fminf(lo[1], c.y); lo[2] = fminf(lo[2], c.z);
        hi[0] = fmaxf(hi[0], c.x); hi[1] = fmaxf(hi[1], c.y); hi[2] = fmaxf(hi[2], c.z);
    }
    __shared__ float s[6][256];
    for (int c = 0; c < 3; ++c) {
        s[c][threadIdx.x] = lo[c];
        s[3 + c][threadIdx.x] = hi[c];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int c = 0; c < 3; ++c) {
                s[c][threadIdx.x] = fminf(s[c][threadIdx.x], s[c][threadIdx.x + w]);
                s[3 + c][threadIdx.x] = fmaxf(s[3 + c][threadIdx.x], s[3 + c][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[blockIdx.x * 8 + threadIdx.x] = s[threadIdx.x][0];
}

// Stage 2: one block folds the partials into (lo.xyz, scale.xyz) for the quantiser.
__global__ void __launch_bounds__(256) k_bounds_final(const float* __restrict__ partial,
                                                       int nparts, float* __restrict__ cb) {
    __shared__ float s[6][256];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = threadIdx.x; i < nparts; i += blockDim.x)
        for (int c = 0; c < 3; ++c) {
            lo[c] = fminf(lo[c], partial[i * 8 + c]);
            hi[c] = fmaxf(hi[c], partial[i * 8 + 3 + c]);
        }
    for (int c = 0; c < 3; ++c) {
        s[c][threadIdx.x] = lo[c];
        s[3 + c][threadIdx.x] = hi[c];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int c = 0; c < 3; ++c) {
                s[c][threadIdx.x] = fminf(s[c][threadIdx.x], s[c][threadIdx.x + w]);
                s[3 + c][threadIdx.x] = fmaxf(s[3 + c][threadIdx.x], s[3 + c][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        const float l = s[threadIdx.x][0], h = s[3 + threadIdx.x][0];
        const float ext = h - l;
        cb[threadIdx.x] = l;
        cb[4 + threadIdx.x] = ext > 0.0f ? 1024.0f / ext : 0.0f;
    }
}

RT_DEV uint32_t expand10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

RT_DEV uint32_t quantize10(float c, float lo, float scale) {
    float q = (c - lo) * scale;
    q = fminf(fmaxf(q, 0.0f), 1023.0f);
    return (uint32_t)q;
}

__global__ void __launch_bounds__(256) k_morton(const float4* __restrict__ cent, uint32_t n,
                                                 const float* __restrict__ cb,
                                                 uint32_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 c = cent[i];
    const uint32_t qx = quantize10(c.x, cb[0], cb[4]);
    const uint32_t qy = quantize10(c.y, cb[1], cb[5]);
    const uint32_t qz = quantize10(c.z, cb[2], cb[6]);
    keys[i] = (expand10(qx) << 2) | (expand10(qy) << 1) | expand10(qz);
    vals[i] = i;
}

// ---- hierarchy ----


RT_DEV int delta(const uint32_t* __restrict__ k, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = k[i], b = k[j];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz((uint32_t)i ^ (uint32_t)j);
}

__global__ void __launch_bounds__(256) k_hierarchy(const uint32_t* __restrict__ keys, uint32_t n,
                                                    float4* __restrict__ nodes,
                                                    int32_t* __restrict__ parent,
                                                    int32_t* __restrict__ rangej,
                                                    int32_t* __restrict__ split,
                                                    uint32_t* __restrict__ top_count,
                                                    int32_t* __restrict__ top_list) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n - 1) return;
    const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int64_t lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int64_t s = 0, t = l;
    for (;;) {
        t = (t + 1) / 2;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    // binary radix tree (used by the refit): children gamma / gamma+1, leaves when their range is
    // a single slot
    const int32_t left = (lo == gamma) ? ~(int32_t)gamma : (int32_t)gamma;
    const int32_t right = (hi == gamma + 1) ? ~(int32_t)(gamma + 1) : (int32_t)(gamma + 1);
    // traversal refs: a child whose range holds <= kLeafMax slots becomes one leaf of that range
    const int64_t nl = gamma - lo + 1, nr = hi - gamma;
    const int32_t tl = nl <= kLeafMax ? leaf_ref((uint32_t)lo, (uint32_t)nl) : (int32_t)gamma;
    const int32_t tr = nr <= kLeafMax ? leaf_ref((uint32_t)(gamma + 1), (uint32_t)nr) : (int32_t)(gamma + 1);
    nodes[(uint64_t)i * 4 + 3] = make_float4(__int_as_float(tl), __int_as_float(tr), 0.0f, 0.0f);
    const uint64_t ni = (uint64_t)n - 1;
    parent[left >= 0 ? (uint64_t)left : ni + (uint64_t)(~left)] = (int32_t)i;
    parent[right >= 0 ? (uint64_t)right : ni + (uint64_t)(~right)] = (int32_t)i;
    rangej[i] = (int32_t)j;
    split[i] = (int32_t)gamma;
    if (i == 0) parent[0] = -1;
    if (lo / kRefitBlock != hi / kRefitBlock) top_list[atomicAdd(top_count, 1u)] = (int32_t)i;
}

// One thread per float4 of the outputs: thread x copies shade float4 x (record x/4) and, for
// x < 3n, tris float4 x (record x/3), so each store instruction writes consecutive 16 B pieces
// and each record's source is read as consecutive pieces by neighbouring lanes.
__global__ void __launch_bounds__(256) k_emit(const uint32_t* __restrict__ order, uint32_t n,
                                              const float4* __restrict__ tri_u,
                                              const float4* __restrict__ shade_u,
                                              float4* __restrict__ tris,
                                              float4* __restrict__ shade) {
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= 4 * (uint64_t)n) return;
    const uint64_t ps = order[x >> 2];
    shade[x] = shade_u[4 * ps + (x & 3)];
    if (x < 3 * (uint64_t)n) {
        const uint64_t s = x / 3;
        tris[x] = tri_u[3 * (uint64_t)order[s] + (x - 3 * s)];
    }
}

// ---- refit ----
// Internal node i covers the sorted leaves [min(i,j), max(i,j)] (j = rangej[i]). With blocks of
// kRefitBlock consecutive leaves, a node whose range lies in one block is "local": its whole
// subtree is refit by that block's workgroup, which synchronises only through LDS. The few nodes
// whose ranges span blocks ("top" nodes) are refit afterwards with device-scope counters.

// Children of internal node i in the binary radix tree: internal ids, or ~slot for single leaves.
RT_DEV void binary_children(const int32_t* __restrict__ rangej, const int32_t* __restrict__ split,
                            int32_t i, int32_t& l, int32_t& r) {
    const int32_t j = rangej[i], g = split[i];
    const int32_t lo = i < j ? i : j, hi = i < j ? j : i;
    l = lo == g ? ~g : g;
    r = hi == g + 1 ? ~(g + 1) : g + 1;
}

RT_DEV void leaf_box(const float4* __restrict__ leafbox, const uint32_t* __restrict__ order,
                     int32_t ref, float* lo, float* hi) {
    const uint64_t p = order[~ref];
    const float4 a = leafbox[2 * p], b = leafbox[2 * p + 1];
    lo[0] = a.x; lo[1] = a.y; lo[2] = a.z;
    hi[0] = b.x; hi[1] = b.y; hi[2] = b.z;
}

// Writes the child boxes into node record `node`, its own box (union) to nodebox, and returns it.
RT_DEV void store_node(float4* __restrict__ nodes, float4* __restrict__ nodebox,
                       float* __restrict__ bounds, int32_t node, const float* l_lo, const float* l_hi,
                       const float* r_lo, const float* r_hi, float* u_lo, float* u_hi) {
    float4* nd = nodes + (uint64_t)node * 4;
    nd[0] = make_float4(l_lo[0], l_hi[0], l_lo[1], l_hi[1]);
    nd[1] = make_float4(r_lo[0], r_hi[0], r_lo[1], r_hi[1]);
    nd[2] = make_float4(l_lo[2], l_hi[2], r_lo[2], r_hi[2]);
    for (int c = 0; c < 3; ++c) {
        u_lo[c] = fminf(l_lo[c], r_lo[c]);
        u_hi[c] = fmaxf(l_hi[c], r_hi[c]);
    }
    nodebox[2 * (uint64_t)node] = make_float4(u_lo[0], u_lo[1], u_lo[2], 0.0f);
    nodebox[2 * (uint64_t)node + 1] = make_float4(u_hi[0], u_hi[1], u_hi[2], 0.0f);
    if (node == 0) {
        bounds[0] = u_lo[0]; bounds[1] = u_lo[1]; bounds[2] = u_lo[2]; bounds[3] = 0.0f;
        bounds[4] = u_hi[0]; bounds[5] = u_hi[1]; bounds[6] = u_hi[2]; bounds[7] = 0.0f;
    }
}

// Range unions for the top refit, per refit block: the prefix and suffix unions of its leaf boxes
// in slot order and the block's total, level 0 of the sparse table. A workgroup scan over the
// block's slots, kRefitBlock / 256 per thread (thread t holds slots s0 .. s0 + kPer - 1 in lo/hi,
// empty boxes past n); tlo/thi are 256-entry LDS rows. Called by all threads of the workgroup.
constexpr int kPer = kRefitBlock / 256;
RT_DEV void block_unions(const float (&lo)[kPer][3], const float (&hi)[kPer][3], uint64_t s0, uint32_t n,
                         float4* __restrict__ pre, float4* __restrict__ suf,
                         float4* __restrict__ sparse, float (*tlo)[256], float (*thi)[256]) {
    const uint32_t t = threadIdx.x;
    // exclusive prefix union of the thread totals, then inclusive per slot
    float tl[3], th[3];
    for (int c = 0; c < 3; ++c) {
        tl[c] = INFINITY;
        th[c] = -INFINITY;
        for (int k = 0; k < kPer; ++k) {
            tl[c] = fminf(tl[c], lo[k][c]);
            th[c] = fmaxf(th[c], hi[k][c]);
        }
        tlo[c][t] = tl[c];
        thi[c][t] = th[c];
    }
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {  // inclusive scan of the totals
        float xl[3], xh[3];
        for (int c = 0; c < 3; ++c) {
            xl[c] = t >= (uint32_t)off ? tlo[c][t - off] : INFINITY;
            xh[c] = t >= (uint32_t)off ? thi[c][t - off] : -INFINITY;
        }
        __syncthreads();
        for (int c = 0; c < 3; ++c) {
            tlo[c][t] = fminf(tlo[c][t], xl[c]);
            thi[c][t] = fmaxf(thi[c][t], xh[c]);
        }
        __syncthreads();
    }
    float rl[3], rh[3];
    for (int c = 0; c < 3; ++c) {
        rl[c] = t > 0 ? tlo[c][t - 1] : INFINITY;
        rh[c] = t > 0 ? thi[c][t - 1] : -INFINITY;
    }
    if (t == 255) {
        sparse[2 * (uint64_t)blockIdx.x] = make_float4(tlo[0][255], tlo[1][255], tlo[2][255], 0.0f);
        sparse[2 * (uint64_t)blockIdx.x + 1] = make_float4(thi[0][255], thi[1][255], thi[2][255], 0.0f);
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        for (int c = 0; c < 3; ++c) {
            rl[c] = fminf(rl[c], lo[k][c]);
            rh[c] = fmaxf(rh[c], hi[k][c]);
        }
        if (s0 + k < n) {
            pre[2 * (s0 + k)] = make_float4(rl[0], rl[1], rl[2], 0.0f);
            pre[2 * (s0 + k) + 1] = make_float4(rh[0], rh[1], rh[2], 0.0f);
        }
    }
    __syncthreads();
    // suffix: the same over the reversed thread order
    for (int c = 0; c < 3; ++c) {
        tlo[c][t] = tl[c];
        thi[c][t] = th[c];
    }
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        float xl[3], xh[3];
        for (int c = 0; c < 3; ++c) {
            xl[c] = t + off < 256u ? tlo[c][t + off] : INFINITY;
            xh[c] = t + off < 256u ? thi[c][t + off] : -INFINITY;
        }
        __syncthreads();
        for (int c = 0; c < 3; ++c) {
            tlo[c][t] = fminf(tlo[c][t], xl[c]);
            thi[c][t] = fmaxf(thi[c][t], xh[c]);
        }
        __syncthreads();
    }
    for (int c = 0; c < 3; ++c) {
        rl[c] = t < 255 ? tlo[c][t + 1] : INFINITY;
        rh[c] = t < 255 ? thi[c][t + 1] : -INFINITY;
    }
#pragma unroll
    for (int k = kPer - 1; k >= 0; --k) {
        for (int c = 0; c < 3; ++c) {
            rl[c] = fminf(rl[c], lo[k][c]);
            rh[c] = fmaxf(rh[c], hi[k][c]);
        }
        if (s0 + k < n) {
            suf[2 * (s0 + k)] = make_float4(rl[0], rl[1], rl[2], 0.0f);
            suf[2 * (s0 + k) + 1] = make_float4(rh[0], rh[1], rh[2], 0.0f);
        }
    }
}


// Phase 1: block b walks up from each of its leaves through the local nodes (those whose leaf range
// lies inside the block; a node's id is an end of its range, so local nodes have ids in the block
// too). The block first copies what the walks read into LDS: the rangej / split / parent entries of
// its node ids and the boxes of its leaves (gathered through `order`). The first thread to reach a
// node stops; the second (LDS counter, workgroup acq_rel) unions the children's boxes from LDS and
// writes the node. Before the walks, the block scans its leaf boxes into range unions.
__global__ void __launch_bounds__(256) k_refit_local(const uint32_t* __restrict__ order, uint32_t n,
                                                      const float4* __restrict__ leafbox,
                                                      const int32_t* __restrict__ parent,
                                                      const int32_t* __restrict__ rangej,
                                                      const int32_t* __restrict__ split,
                                                      float4* __restrict__ nodes,
                                                      float4* __restrict__ nodebox,
                                                      float* __restrict__ bounds,
                                                      float4* __restrict__ pre,
                                                      float4* __restrict__ suf,
                                                      float4* __restrict__ sparse) {
    __shared__ uint32_t lflag[kRefitBlock];
    __shared__ float tlo[3][256], thi[3][256];
    __shared__ float lbox[kRefitBlock][6];   // boxes of the block's internal nodes
    __shared__ float lleaf[kRefitBlock][6];  // boxes of the block's leaves (slots)
    __shared__ int32_t lj[kRefitBlock], lsplit[kRefitBlock], lpar[kRefitBlock];
    const int64_t blo = (int64_t)blockIdx.x * kRefitBlock;
    const int64_t bhi = std::min<int64_t>(blo + kRefitBlock, n);  // exclusive
    const int64_t ni = (int64_t)n - 1;
    for (int k = threadIdx.x; k < kRefitBlock; k += blockDim.x) {
        lflag[k] = 0;
        if (blo + k < ni) {
            lj[k] = rangej[blo + k];
            lsplit[k] = split[blo + k];
            lpar[k] = parent[blo + k];
        }
        if (blo + k < bhi) {
            const uint64_t p = order[blo + k];
            const float4 a = leafbox[2 * p], b = leafbox[2 * p + 1];
            lleaf[k][0] = a.x; lleaf[k][1] = a.y; lleaf[k][2] = a.z;
            lleaf[k][3] = b.x; lleaf[k][4] = b.y; lleaf[k][5] = b.z;
        }
    }
    __syncthreads();
    {  // the block's range unions, from the leaf boxes already in LDS
        float lo[kPer][3], hi[kPer][3];
        const uint32_t l0 = threadIdx.x * kPer;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const bool in = blo + l0 + k < bhi;
            for (int c = 0; c < 3; ++c) {
                lo[k][c] = in ? lleaf[l0 + k][c] : INFINITY;
                hi[k][c] = in ? lleaf[l0 + k][3 + c] : -INFINITY;
            }
        }
        block_unions(lo, hi, (uint64_t)blo + l0, n, pre, suf, sparse, tlo, thi);
    }
    for (int64_t s = blo + threadIdx.x; s < bhi; s += blockDim.x) {
        int32_t node = parent[ni + s];
        while (node >= blo && node < bhi && node < ni) {
            const int32_t li = (int32_t)(node - blo);
            const int64_t j = lj[li];
            const int64_t lo = std::min<int64_t>(node, j), hi = std::max<int64_t>(node, j);
            if (lo < blo || hi >= bhi) break;  // a top node: phase 2
            const uint32_t prev = __hip_atomic_fetch_add(&lflag[li], 1u, __ATOMIC_ACQ_REL,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP);
            if (prev == 0) break;
            const int32_t g = lsplit[li];
            const int32_t l = lo == g ? ~g : g, r = hi == g + 1 ? ~(g + 1) : g + 1;
            const float* lb = l < 0 ? lleaf[~l - blo] : lbox[l - blo];
            const float* rb = r < 0 ? lleaf[~r - blo] : lbox[r - blo];
            float l_lo[3], l_hi[3], r_lo[3], r_hi[3], u_lo[3], u_hi[3];
            for (int c = 0; c < 3; ++c) {
                l_lo[c] = lb[c];
                l_hi[c] = lb[3 + c];
                r_lo[c] = rb[c];
                r_hi[c] = rb[3 + c];
            }
            store_node(nodes, nodebox, bounds, node, l_lo, l_hi, r_lo, r_hi, u_lo, u_hi);
            for (int c = 0; c < 3; ++c) {
                lbox[li][c] = u_lo[c];
                lbox[li][3 + c] = u_hi[c];
            }
            node = lpar[li];
        }
    }
}

RT_DEV bool is_local(const int32_t* __restrict__ rangej, int32_t ref) {
    if (ref < 0) return true;
    const int32_t j = rangej[ref];
    const int32_t lo = ref < j ? ref : j, hi = ref < j ? j : ref;
    return lo / kRefitBlock == hi / kRefitBlock;
}

RT_DEV void node_box(const float4* __restrict__ leafbox, const uint32_t* __restrict__ order,
                     const float4* nodebox, int32_t ref, float* lo, float* hi) {
    if (ref < 0) {
        leaf_box(leafbox, order, ref, lo, hi);
    } else {
        const float4 a = nodebox[2 * (uint64_t)ref], b = nodebox[2 * (uint64_t)ref + 1];
        lo[0] = a.x; lo[1] = a.y; lo[2] = a.z;
        hi[0] = b.x; hi[1] = b.y; hi[2] = b.z;
    }
}

// Sparse table levels 1.. in one workgroup: level k, block b = union of levels k-1 at b and
// b + 2^(k-1) (the workgroup's barrier orders the levels' global stores and loads).
__global__ void __launch_bounds__(1024) k_block_sparse(float4* __restrict__ sparse, uint32_t nb,
                                                        uint32_t levels) {
    for (uint32_t k = 1; k < levels; ++k) {
        const uint32_t half = 1u << (k - 1), cnt = nb - (2u * half) + 1u;
        const float4* src = sparse + 2 * (uint64_t)(k - 1) * nb;
        float4* dst = sparse + 2 * (uint64_t)k * nb;
        for (uint32_t b = threadIdx.x; b < cnt; b += blockDim.x) {
            const float4 a0 = src[2 * (uint64_t)b], a1 = src[2 * (uint64_t)b + 1];
            const float4 b0 = src[2 * (uint64_t)(b + half)], b1 = src[2 * (uint64_t)(b + half) + 1];
            dst[2 * (uint64_t)b] = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), 0.0f);
            dst[2 * (uint64_t)b + 1] = make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), 0.0f);
        }
        __syncthreads();
    }
}

// Union of the leaf boxes of slots [lo, hi], which span at least two refit blocks: the suffix of
// lo's block, the prefix of hi's block and the blocks between (two overlapping sparse-table
// entries; min/max are exact, so overlap and order do not change the result).
RT_DEV void range_box(const float4* __restrict__ pre, const float4* __restrict__ suf,
                      const float4* __restrict__ sparse, uint32_t nb, int32_t lo, int32_t hi,
                      float* blo, float* bhi) {
    const float4 s0 = suf[2 * (uint64_t)lo], s1 = suf[2 * (uint64_t)lo + 1];
    const float4 p0 = pre[2 * (uint64_t)hi], p1 = pre[2 * (uint64_t)hi + 1];
    blo[0] = fminf(s0.x, p0.x); blo[1] = fminf(s0.y, p0.y); blo[2] = fminf(s0.z, p0.z);
    bhi[0] = fmaxf(s1.x, p1.x); bhi[1] = fmaxf(s1.y, p1.y); bhi[2] = fmaxf(s1.z, p1.z);
    const int32_t b0 = lo / kRefitBlock + 1, b1 = hi / kRefitBlock - 1;
    if (b0 <= b1) {
        const uint32_t len = (uint32_t)(b1 - b0 + 1);
        const uint32_t k = 31u - __builtin_clz(len);
        const float4* lv = sparse + 2 * (uint64_t)k * nb;
        const uint64_t x = (uint64_t)b0, y = (uint64_t)(b1 - (int32_t)(1u << k) + 1);
        const float4 a0 = lv[2 * x], a1 = lv[2 * x + 1], c0 = lv[2 * y], c1 = lv[2 * y + 1];
        blo[0] = fminf(blo[0], fminf(a0.x, c0.x));
        blo[1] = fminf(blo[1], fminf(a0.y, c0.y));
        blo[2] = fminf(blo[2], fminf(a0.z, c0.z));
        bhi[0] = fmaxf(bhi[0], fmaxf(a1.x, c1.x));
        bhi[1] = fmaxf(bhi[1], fmaxf(a1.y, c1.y));
        bhi[2] = fmaxf(bhi[2], fmaxf(a1.z, c1.z));
    }
}

// Phase 3, one thread per top node (the list written by k_hierarchy): each child's box is its
// range union (a top child), or the box phase 1 computed (a local child or a leaf). No node waits on
// another.
__global__ void __launch_bounds__(256) k_refit_top(const uint32_t* __restrict__ order, uint32_t n,
                                                    const float4* __restrict__ leafbox,
                                                    const int32_t* __restrict__ rangej,
                                                    const int32_t* __restrict__ split,
                                                    float4* __restrict__ nodes, float4* nodebox,
                                                    float* __restrict__ bounds,
                                                    const int32_t* __restrict__ top_list,
                                                    const uint32_t* __restrict__ top_count,
                                                    const float4* __restrict__ pre,
                                                    const float4* __restrict__ suf,
                                                    const float4* __restrict__ sparse, uint32_t nb) {
    const uint32_t cnt = *top_count;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < cnt; t += gridDim.x * blockDim.x) {
        const int32_t node = top_list[t];
        int32_t l, r;
        binary_children(rangej, split, node, l, r);
        float l_lo[3], l_hi[3], r_lo[3], r_hi[3], u_lo[3], u_hi[3];
        if (is_local(rangej, l)) {
            node_box(leafbox, order, nodebox, l, l_lo, l_hi);
        } else {
            const int32_t j = rangej[l];
            range_box(pre, suf, sparse, nb, l < j ? l : j, l < j ? j : l, l_lo, l_hi);
        }
        if (is_local(rangej, r)) {
            node_box(leafbox, order, nodebox, r, r_lo, r_hi);
        } else {
            const int32_t j = rangej[r];
            range_box(pre, suf, sparse, nb, r < j ? r : j, r < j ? j : r, r_lo, r_hi);
        }
        store_node(nodes, nodebox, bounds, node, l_lo, l_hi, r_lo, r_hi, u_lo, u_hi);
    }
}

__global__ void k_zero_word(uint32_t* w) { *w = 0u; }

// Traversals keep pending nodes on stacks sized for trees at most kMaxTreeDepth internal levels
// deep (the Karras tree over 62-bit keys has at most 62). A PLOC tree has no such bound (a chain of
// clusters can nest deeper), so its depth is measured: per triangle, the internal nodes above it
// (leaves of the parent array at ni + primitive), counted up to kMaxTreeDepth + 1.
constexpr uint32_t kMaxTreeDepth = 62;

__global__ void __launch_bounds__(256) k_ploc_depth(uint32_t n, const int32_t* __restrict__ parent,
                                                     uint32_t* __restrict__ depth) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t x = parent[(uint64_t)(n - 1) + i];
    uint32_t d = 0;
    while (x >= 0 && d <= kMaxTreeDepth) {
        ++d;
        x = parent[x];
    }
    if (d > kMaxTreeDepth) atomicMax(depth, d);
}

// ---- PLOC builder (parallel locally-ordered clustering, Meister & Bittner 2018) ----
// Clusters start as the triangles in Morton order (positions 0..n-1, refs ~prim, padded leaf
// boxes). Each iteration: every cluster i picks its nearest neighbour NN(i) among positions
// [i-R, i+R] (smallest surface area of the union box, ties to the smaller position); mutual pairs
// merge into a new node kept at the lower position, the higher one drops out, survivors keep their
// order. An iteration with no mutual pair (possible only with NaN boxes) merges positions 0 and 1.
// The k-th node created (over iterations, in position order) gets id n-2-k: the root is 0 and
// children have larger ids than their parents. Large iterations run as four grid launches
// (neighbours; per-chunk counts; chunk offsets and the next cluster count; compaction); once at
// most kPlocTail clusters remain, one workgroup finishes in a single launch. Subtrees of at most
// kLeafMax triangles whose parent is larger become leaves: leaf roots take consecutive slot ranges
// in the order of their smallest Morton position, their triangles left to right.
constexpr int kPlocRadius = 16;
constexpr int kPlocTile = 256;        // clusters per neighbour-search tile (one per thread)
constexpr int kPlocChunk = 1024;      // clusters per count / compaction chunk (4 per thread)
constexpr int kPlocGrid = 1024;       // blocks of the grid-stride iteration kernels
constexpr int32_t kPlocTail = 4096;   // cluster count handed to the one-workgroup tail

// state[4 * parity]: clusters m, nodes created, forced merge, buffer parity of the clusters
struct PlocState {
    int32_t m, created, forced, buf;
};

RT_DEV float union_area(const float4& alo, const float4& ahi, const float4& blo, const float4& bhi) {
    const float dx = fmaxf(ahi.x, bhi.x) - fminf(alo.x, blo.x);
    const float dy = fmaxf(ahi.y, bhi.y) - fminf(alo.y, blo.y);
    const float dz = fmaxf(ahi.z, bhi.z) - fminf(alo.z, blo.z);
    return dx * dy + dy * dz + dz * dx;
}

RT_DEV int32_t ploc_size(const int32_t* __restrict__ nsize, int32_t c) { return c >= 0 ? nsize[c] : 1; }

__global__ void __launch_bounds__(256) k_ploc_init(const uint32_t* __restrict__ morder, uint32_t n,
                                                    const float4* __restrict__ leafbox,
                                                    int32_t* __restrict__ cid, float4* __restrict__ cbox,
                                                    uint32_t* __restrict__ pos_of,
                                                    int32_t* __restrict__ parent, int32_t* __restrict__ state) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = morder[i];
    pos_of[p] = i;
    cid[i] = ~(int32_t)p;
    cbox[2 * (uint64_t)i] = leafbox[2 * (uint64_t)p];
    cbox[2 * (uint64_t)i + 1] = leafbox[2 * (uint64_t)p + 1];
    if (i == 0) {
        parent[0] = -1;
        state[0] = (int32_t)n;
        state[1] = 0;
        state[2] = 0;
        state[3] = 0;
    }
}

// nearest neighbour of position i among [i-R, i+R] (m >= 2), reading boxes through `box(j)`
template <typename BoxAt>
RT_DEV int32_t ploc_nn(int32_t i, int32_t m, BoxAt box) {
    const int32_t lo = i - kPlocRadius < 0 ? 0 : i - kPlocRadius;
    const int32_t hi = i + kPlocRadius > m - 1 ? m - 1 : i + kPlocRadius;
    float4 ilo, ihi, jlo, jhi;
    box(i, ilo, ihi);
    int32_t bj = lo == i ? i + 1 : lo;
    box(bj, jlo, jhi);
    float best = union_area(ilo, ihi, jlo, jhi);
    for (int32_t j = bj + 1; j <= hi; ++j) {
        if (j == i) continue;
        box(j, jlo, jhi);
        const float a = union_area(ilo, ihi, jlo, jhi);
        if (a < best) {
            best = a;
            bj = j;
        }
    }
    return bj;
}

__global__ void __launch_bounds__(kPlocTile) k_ploc_nn(int it, const int32_t* __restrict__ state,
                                                        const float4* __restrict__ cbox0,
                                                        const float4* __restrict__ cbox1,
                                                        int32_t* __restrict__ nn) {
    const int32_t m = state[4 * (it & 1)];
    if (m <= kPlocTail) return;
    const float4* __restrict__ B = (it & 1) ? cbox1 : cbox0;
    __shared__ float4 wlo[kPlocTile + 2 * kPlocRadius], whi[kPlocTile + 2 * kPlocRadius];
    for (int32_t base = blockIdx.x * kPlocTile; base < m; base += gridDim.x * kPlocTile) {
        const int32_t wb = base - kPlocRadius;
        for (int k = threadIdx.x; k < kPlocTile + 2 * kPlocRadius; k += kPlocTile) {
            const int32_t g = wb + k;
            if (g >= 0 && g < m) {
                wlo[k] = B[2 * (uint64_t)g];
                whi[k] = B[2 * (uint64_t)g + 1];
            }
        }
        __syncthreads();
        const int32_t i = base + (int32_t)threadIdx.x;
        if (i < m)
            nn[i] = ploc_nn(i, m, [&](int32_t j, float4& lo, float4& hi) {
                lo = wlo[j - wb];
                hi = whi[j - wb];
            });
        __syncthreads();
    }
}

// keep / merge flags of position i (forced: the iteration's fallback merge of 0 and 1)
RT_DEV void ploc_flags(const int32_t* __restrict__ nn, int32_t i, bool forced, uint32_t& keep,
                       uint32_t& merge) {
    if (forced) {
        keep = i != 1;
        merge = i == 0;
        return;
    }
    const int32_t j = nn[i];
    const bool mutual = nn[j] == i;
    keep = !mutual || i < j;
    merge = mutual && i < j;
}

// Exclusive scan of one value per thread over a block of T threads (T = blockDim.x); `total` is
// the block's sum. The PLOC kernels scan (keep, merge) counts packed as keep | merge << 32.
template <int T>
RT_DEV uint64_t block_exclusive_scan(uint64_t v, uint64_t* lds, uint64_t& total) {
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int d = 1; d < T; d <<= 1) {
        const uint64_t x = t >= d ? lds[t - d] : 0;
        __syncthreads();
        lds[t] += x;
        __syncthreads();
    }
    total = lds[T - 1];
    const uint64_t incl = lds[t];
    __syncthreads();
    return incl - v;
}

// per chunk of kPlocChunk positions: number kept and number merged
__global__ void __launch_bounds__(256) k_ploc_count(int it, const int32_t* __restrict__ state,
                                                     const int32_t* __restrict__ nn,
                                                     uint32_t* __restrict__ part) {
    const int32_t m = state[4 * (it & 1)];
    if (m <= kPlocTail) return;
    __shared__ uint64_t lds[256];
    const int32_t nchunk = (m + kPlocChunk - 1) / kPlocChunk;
    for (int32_t c = blockIdx.x; c < nchunk; c += gridDim.x) {
        uint64_t v = 0;
        for (int q = 0; q < 4; ++q) {
            const int32_t i = c * kPlocChunk + (int32_t)threadIdx.x * 4 + q;
            if (i < m) {
                uint32_t keep, merge;
                ploc_flags(nn, i, false, keep, merge);
                v += (uint64_t)keep | ((uint64_t)merge << 32);
            }
        }
        uint64_t total;
        block_exclusive_scan<256>(v, lds, total);
        if (threadIdx.x == 0) {
            part[2 * (uint64_t)c] = (uint32_t)total;
            part[2 * (uint64_t)c + 1] = (uint32_t)(total >> 32);
        }
    }
}

// one workgroup: chunk counts -> exclusive offsets (in place); the next iteration's state
__global__ void __launch_bounds__(1024) k_ploc_offsets(int it, int32_t* __restrict__ state,
                                                        uint32_t* __restrict__ part) {
    const int32_t* cur = state + 4 * (it & 1);
    int32_t* nxt = state + 4 * ((it + 1) & 1);
    const int32_t m = cur[0], created = cur[1];
    if (m <= kPlocTail) {
        if (threadIdx.x == 0) {
            nxt[0] = m;
            nxt[1] = created;
            nxt[2] = 0;
            nxt[3] = cur[3];
        }
        return;
    }
    __shared__ uint64_t lds[1024];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int32_t nchunk = (m + kPlocChunk - 1) / kPlocChunk;
    for (int32_t b = 0; b < nchunk; b += 1024) {
        const int32_t c = b + (int32_t)threadIdx.x;
        const uint64_t v = c < nchunk ? (uint64_t)part[2 * (uint64_t)c] | ((uint64_t)part[2 * (uint64_t)c + 1] << 32) : 0;
        uint64_t total;
        const uint64_t ex = block_exclusive_scan<1024>(v, lds, total) + carry;
        if (c < nchunk) {
            part[2 * (uint64_t)c] = (uint32_t)ex;
            part[2 * (uint64_t)c + 1] = (uint32_t)(ex >> 32);
        }
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const uint32_t kept = (uint32_t)carry, merged = (uint32_t)(carry >> 32);
        const bool forced = merged == 0;
        nxt[0] = forced ? m - 1 : (int32_t)kept;
        nxt[1] = created + (forced ? 1 : (int32_t)merged);
        nxt[2] = forced ? 1 : 0;
        nxt[3] = (it + 1) & 1;
    }
}

// writes the merged node `id` from the clusters at positions i < j
RT_DEV void ploc_merge(uint32_t n, int32_t id, int32_t ci, int32_t cj, float4 ilo, float4 ihi,
                       float4 jlo, float4 jhi, int32_t* __restrict__ nleft, int32_t* __restrict__ nright,
                       int32_t* __restrict__ nsize, int32_t* __restrict__ parent, float4* __restrict__ nodes,
                       float* __restrict__ bounds, float4& ulo, float4& uhi) {
    const uint64_t ni = (uint64_t)n - 1;
    nleft[id] = ci;
    nright[id] = cj;
    nsize[id] = ploc_size(nsize, ci) + ploc_size(nsize, cj);
    parent[ci >= 0 ? (uint64_t)ci : ni + (uint64_t)(~ci)] = id;
    parent[cj >= 0 ? (uint64_t)cj : ni + (uint64_t)(~cj)] = id;
    float4* nd = nodes + (uint64_t)id * 4;
    nd[0] = make_float4(ilo.x, ihi.x, ilo.y, ihi.y);
    nd[1] = make_float4(jlo.x, jhi.x, jlo.y, jhi.y);
    nd[2] = make_float4(ilo.z, ihi.z, jlo.z, jhi.z);
    ulo = make_float4(fminf(ilo.x, jlo.x), fminf(ilo.y, jlo.y), fminf(ilo.z, jlo.z), 0.0f);
    uhi = make_float4(fmaxf(ihi.x, jhi.x), fmaxf(ihi.y, jhi.y), fmaxf(ihi.z, jhi.z), 0.0f);
    if (id == 0) {
        bounds[0] = ulo.x; bounds[1] = ulo.y; bounds[2] = ulo.z; bounds[3] = 0.0f;
        bounds[4] = uhi.x; bounds[5] = uhi.y; bounds[6] = uhi.z; bounds[7] = 0.0f;
    }
}

// compaction of one chunk: survivors to their new positions, merged pairs as new nodes
__global__ void __launch_bounds__(256) k_ploc_compact(int it, uint32_t n, const int32_t* __restrict__ state,
                                                       const int32_t* __restrict__ nn,
                                                       const uint32_t* __restrict__ part,
                                                       int32_t* __restrict__ cid0, int32_t* __restrict__ cid1,
                                                       float4* __restrict__ cbox0, float4* __restrict__ cbox1,
                                                       int32_t* __restrict__ nleft, int32_t* __restrict__ nright,
                                                       int32_t* __restrict__ nsize, int32_t* __restrict__ parent,
                                                       float4* __restrict__ nodes, float* __restrict__ bounds) {
    const int32_t* cur = state + 4 * (it & 1);
    const int32_t m = cur[0], created = cur[1];
    if (m <= kPlocTail) return;
    const bool forced = state[4 * ((it + 1) & 1) + 2] != 0;
    const int32_t* __restrict__ C = (it & 1) ? cid1 : cid0;
    const float4* __restrict__ B = (it & 1) ? cbox1 : cbox0;
    int32_t* __restrict__ C2 = (it & 1) ? cid0 : cid1;
    float4* __restrict__ B2 = (it & 1) ? cbox0 : cbox1;
    __shared__ uint64_t lds[256];
    const int32_t nchunk = (m + kPlocChunk - 1) / kPlocChunk;
    for (int32_t c = blockIdx.x; c < nchunk; c += gridDim.x) {
        uint32_t keep[4], merge[4];
        uint64_t v = 0;
        for (int q = 0; q < 4; ++q) {
            const int32_t i = c * kPlocChunk + (int32_t)threadIdx.x * 4 + q;
            keep[q] = merge[q] = 0;
            if (i < m) ploc_flags(nn, i, forced, keep[q], merge[q]);
            v += (uint64_t)keep[q] | ((uint64_t)merge[q] << 32);
        }
        uint64_t total;
        uint64_t ex = block_exclusive_scan<256>(v, lds, total);
        uint32_t dst = part[2 * (uint64_t)c] + (uint32_t)ex;
        uint32_t rank = (forced ? 0u : part[2 * (uint64_t)c + 1]) + (uint32_t)(ex >> 32);
        for (int q = 0; q < 4; ++q) {
            const int32_t i = c * kPlocChunk + (int32_t)threadIdx.x * 4 + q;
            if (!keep[q]) continue;
            if (forced) dst = i == 0 ? 0u : (uint32_t)(i - 1);
            const float4 ilo = B[2 * (uint64_t)i], ihi = B[2 * (uint64_t)i + 1];
            if (merge[q]) {
                const int32_t j = forced ? 1 : nn[i];
                const int32_t id = (int32_t)n - 2 - (created + (int32_t)rank);
                float4 ulo, uhi;
                ploc_merge(n, id, C[i], C[j], ilo, ihi, B[2 * (uint64_t)j], B[2 * (uint64_t)j + 1], nleft,
                           nright, nsize, parent, nodes, bounds, ulo, uhi);
                C2[dst] = id;
                B2[2 * (uint64_t)dst] = ulo;
                B2[2 * (uint64_t)dst + 1] = uhi;
                ++rank;
            } else {
                C2[dst] = C[i];
                B2[2 * (uint64_t)dst] = ilo;
                B2[2 * (uint64_t)dst + 1] = ihi;
            }
            ++dst;
        }
    }
}

// The remaining iterations (m <= kPlocTail normally; any m is handled) in one workgroup, from the
// state left in state[4 * slot].
__global__ void __launch_bounds__(1024) k_ploc_tail(int slot, uint32_t n, int32_t* __restrict__ state,
                                                     int32_t* __restrict__ nn, int32_t* __restrict__ cid0,
                                                     int32_t* __restrict__ cid1, float4* __restrict__ cbox0,
                                                     float4* __restrict__ cbox1, int32_t* __restrict__ nleft,
                                                     int32_t* __restrict__ nright, int32_t* __restrict__ nsize,
                                                     int32_t* __restrict__ parent, float4* __restrict__ nodes,
                                                     float* __restrict__ bounds) {
    __shared__ uint64_t lds[1024];
    __shared__ int32_t sm, screated;
    __shared__ uint32_t smerged;
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) {
        sm = state[4 * slot];
        screated = state[4 * slot + 1];
    }
    int par = state[4 * slot + 3];
    __syncthreads();
    for (;;) {
        const int32_t m = sm, created = screated;
        if (m <= 1) break;
        const int32_t* __restrict__ C = par ? cid1 : cid0;
        const float4* __restrict__ B = par ? cbox1 : cbox0;
        int32_t* __restrict__ C2 = par ? cid0 : cid1;
        float4* __restrict__ B2 = par ? cbox0 : cbox1;
        for (int32_t i = threadIdx.x; i < m; i += blockDim.x)
            nn[i] = ploc_nn(i, m, [&](int32_t j, float4& lo, float4& hi) {
                lo = B[2 * (uint64_t)j];
                hi = B[2 * (uint64_t)j + 1];
            });
        if (threadIdx.x == 0) smerged = 0;
        __syncthreads();
        uint32_t mine = 0;
        for (int32_t i = threadIdx.x; i < m; i += blockDim.x) {
            uint32_t keep, merge;
            ploc_flags(nn, i, false, keep, merge);
            mine += merge;
        }
        if (mine) atomicAdd(&smerged, mine);
        __syncthreads();
        const bool forced = smerged == 0;
        if (threadIdx.x == 0) carry = 0;
        __syncthreads();
        for (int32_t b = 0; b < m; b += 4 * (int32_t)blockDim.x) {
            uint32_t keep[4], merge[4];
            uint64_t v = 0;
            for (int q = 0; q < 4; ++q) {
                const int32_t i = b + (int32_t)threadIdx.x * 4 + q;
                keep[q] = merge[q] = 0;
                if (i < m) ploc_flags(nn, i, forced, keep[q], merge[q]);
                v += (uint64_t)keep[q] | ((uint64_t)merge[q] << 32);
            }
            uint64_t total;
            const uint64_t ex = block_exclusive_scan<1024>(v, lds, total) + carry;
            uint32_t dst = (uint32_t)ex, rank = (uint32_t)(ex >> 32);
            for (int q = 0; q < 4; ++q) {
                const int32_t i = b + (int32_t)threadIdx.x * 4 + q;
                if (!keep[q]) continue;
                if (forced) dst = i == 0 ? 0u : (uint32_t)(i - 1);
                const float4 ilo = B[2 * (uint64_t)i], ihi = B[2 * (uint64_t)i + 1];
                if (merge[q]) {
                    const int32_t j = forced ? 1 : nn[i];
                    const int32_t id = (int32_t)n - 2 - (created + (int32_t)rank);
                    float4 ulo, uhi;
                    ploc_merge(n, id, C[i], C[j], ilo, ihi, B[2 * (uint64_t)j], B[2 * (uint64_t)j + 1],
                               nleft, nright, nsize, parent, nodes, bounds, ulo, uhi);
                    C2[dst] = id;
                    B2[2 * (uint64_t)dst] = ulo;
                    B2[2 * (uint64_t)dst + 1] = uhi;
                    ++rank;
                } else {
                    C2[dst] = C[i];
                    B2[2 * (uint64_t)dst] = ilo;
                    B2[2 * (uint64_t)dst + 1] = ihi;
                }
                ++dst;
            }
            __syncthreads();
            if (threadIdx.x == 0) carry += total;
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            sm = forced ? m - 1 : (int32_t)(uint32_t)carry;
            screated = created + (forced ? 1 : (int32_t)(carry >> 32));
        }
        par ^= 1;
        __syncthreads();
    }
}

// leaf root of primitive p: walk up while the parent holds at most kLeafMax triangles
RT_DEV int32_t ploc_leaf_root(uint32_t n, const int32_t* __restrict__ parent,
                              const int32_t* __restrict__ nsize, uint32_t p) {
    const uint64_t ni = (uint64_t)n - 1;
    int32_t x = ~(int32_t)p;
    for (;;) {
        const int32_t q = parent[x >= 0 ? (uint64_t)x : ni + (uint64_t)(~x)];
        if (q < 0 || nsize[q] > kLeafMax) return x;
        x = q;
    }
}

// smallest Morton position among the (at most kLeafMax) triangles under x
RT_DEV uint32_t ploc_minpos(int32_t x, const int32_t* __restrict__ nleft, const int32_t* __restrict__ nright,
                            const uint32_t* __restrict__ pos_of) {
    int32_t stk[2 * kLeafMax];
    int sp = 0;
    stk[sp++] = x;
    uint32_t best = 0xFFFFFFFFu;
    while (sp) {
        const int32_t c = stk[--sp];
        if (c < 0) {
            const uint32_t q = pos_of[~c];
            best = q < best ? q : best;
        } else if (sp + 2 <= 2 * kLeafMax) {  // always true for a subtree of <= kLeafMax triangles
            stk[sp++] = nright[c];
            stk[sp++] = nleft[c];
        }
    }
    return best;
}

__global__ void __launch_bounds__(256) k_ploc_leafroots(uint32_t n, const uint32_t* __restrict__ morder,
                                                         const int32_t* __restrict__ parent,
                                                         const int32_t* __restrict__ nsize,
                                                         const int32_t* __restrict__ nleft,
                                                         const int32_t* __restrict__ nright,
                                                         const uint32_t* __restrict__ pos_of,
                                                         uint32_t* __restrict__ scan) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t x = ploc_leaf_root(n, parent, nsize, morder[i]);
    scan[i] = ploc_minpos(x, nleft, nright, pos_of) == i ? (uint32_t)ploc_size(nsize, x) : 0u;
}

// exclusive scan of scan[0..n) in place: per-chunk totals, chunk offsets, chunk scans
__global__ void __launch_bounds__(256) k_scan_chunks(uint32_t n, const uint32_t* __restrict__ a,
                                                      uint32_t* __restrict__ part) {
    __shared__ uint64_t lds[256];
    const uint32_t c = blockIdx.x;
    uint64_t v = 0;
    for (int q = 0; q < 4; ++q) {
        const uint64_t i = (uint64_t)c * kPlocChunk + threadIdx.x * 4 + q;
        if (i < n) v += a[i];
    }
    uint64_t total;
    block_exclusive_scan<256>(v, lds, total);
    if (threadIdx.x == 0) part[c] = (uint32_t)total;
}

__global__ void __launch_bounds__(1024) k_scan_top(uint32_t nchunk, uint32_t* __restrict__ part) {
    __shared__ uint64_t lds[1024];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < nchunk; b += 1024) {
        const uint32_t c = b + threadIdx.x;
        const uint64_t v = c < nchunk ? part[c] : 0;
        uint64_t total;
        const uint64_t ex = block_exclusive_scan<1024>(v, lds, total) + carry;
        if (c < nchunk) part[c] = (uint32_t)ex;
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_scan_apply(uint32_t n, uint32_t* __restrict__ a,
                                                     const uint32_t* __restrict__ part) {
    __shared__ uint64_t lds[256];
    const uint32_t c = blockIdx.x;
    uint32_t x[4];
    uint64_t v = 0;
    for (int q = 0; q < 4; ++q) {
        const uint64_t i = (uint64_t)c * kPlocChunk + threadIdx.x * 4 + q;
        x[q] = i < n ? a[i] : 0u;
        v += x[q];
    }
    uint64_t total;
    uint64_t run = block_exclusive_scan<256>(v, lds, total) + part[c];
    for (int q = 0; q < 4; ++q) {
        const uint64_t i = (uint64_t)c * kPlocChunk + threadIdx.x * 4 + q;
        if (i < n) a[i] = (uint32_t)run;
        run += x[q];
    }
}

// slots of the leaf root whose smallest Morton position is i: its triangles left to right from
// the root's first slot; first slots of its internal nodes
__global__ void __launch_bounds__(256) k_ploc_slots(uint32_t n, const uint32_t* __restrict__ morder,
                                                     const int32_t* __restrict__ parent,
                                                     const int32_t* __restrict__ nsize,
                                                     const int32_t* __restrict__ nleft,
                                                     const int32_t* __restrict__ nright,
                                                     const uint32_t* __restrict__ pos_of,
                                                     const uint32_t* __restrict__ scan,
                                                     uint32_t* __restrict__ slot_order,
                                                     uint32_t* __restrict__ slot_of,
                                                     uint32_t* __restrict__ fslot) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t x = ploc_leaf_root(n, parent, nsize, morder[i]);
    if (ploc_minpos(x, nleft, nright, pos_of) != i) return;
    uint32_t next = scan[i];
    int32_t stk[2 * kLeafMax];
    int sp = 0;
    stk[sp++] = x;
    while (sp) {
        const int32_t c = stk[--sp];
        if (c < 0) {
            slot_order[next] = (uint32_t)~c;
            slot_of[~c] = next;
            ++next;
        } else if (sp + 2 <= 2 * kLeafMax) {  // always true for a subtree of <= kLeafMax triangles
            fslot[c] = next;
            stk[sp++] = nright[c];
            stk[sp++] = nleft[c];
        }
    }
}

// traversal refs of every internal node: a child of more than kLeafMax triangles by id, else the
// leaf of its slots
__global__ void __launch_bounds__(256) k_ploc_refs(uint32_t n, const int32_t* __restrict__ nleft,
                                                    const int32_t* __restrict__ nright,
                                                    const int32_t* __restrict__ nsize,
                                                    const uint32_t* __restrict__ slot_of,
                                                    const uint32_t* __restrict__ fslot,
                                                    float4* __restrict__ nodes) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n - 1) return;
    auto tref = [&](int32_t c) -> int32_t {
        if (c < 0) return leaf_ref(slot_of[~c], 1u);
        if (nsize[c] > kLeafMax) return c;
        return leaf_ref(fslot[c], (uint32_t)nsize[c]);
    };
    nodes[(uint64_t)x * 4 + 3] = make_float4(__int_as_float(tref(nleft[x])), __int_as_float(tref(nright[x])),
                                             0.0f, 0.0f);
}

// big PLOC iterations to launch for n clusters: enough for m to fall to kPlocTail at 0.88 per
// iteration (measured ~0.83-0.90), plus margin; the tail finishes whatever remains
inline int ploc_iterations(uint32_t n) {
    if (n <= (uint32_t)kPlocTail) return 0;
    return (int)std::ceil(std::log((double)n / kPlocTail) / std::log(1.0 / 0.88)) + 4;
}


// ---- 4-wide nodes ----
// nodes4[i] (8 x float4, 128 B) lists the grandchildren of binary node i reached through its
// internal children, and its leaf children directly, in the order (left..., right...):
//   [0]=lo.x[4] [1]=hi.x[4] [2]=lo.y[4] [3]=hi.y[4] [4]=lo.z[4] [5]=hi.z[4] [6]=refs[4] [7]=0
// Unused slots hold ref RT_EMPTY_ROOT and an empty box (lo=+inf, hi=-inf). Every binary node gets
// a record (traversal from the root only reaches every other level), so refs stay binary ids.
// 4-wide record of binary node i: the boxes and references of its grandchildren (a child that is a
// leaf stands for itself), empty slots padded. The block's 256 records are staged in LDS (a
// 144-byte stride keeps the per-lane writes off one bank) and stored as consecutive float4s, so
// every store instruction writes 4 KB of the output contiguously. (Loading the block's binary
// records through LDS as well measured slower: they are read once and the child reads hit L2.)
__global__ void __launch_bounds__(256) k_wide(uint32_t n, const float4* __restrict__ nodes,
                                              float4* __restrict__ nodes4) {
    constexpr int kPitch = 9;  // float4s per staged record
    __shared__ float4 stage[256 * kPitch];
    const uint64_t base = (uint64_t)blockIdx.x * blockDim.x;
    const uint64_t i = base + threadIdx.x;
    const uint64_t m = (uint64_t)n - 1;  // internal nodes
    const uint64_t rows = m - base < 256 ? m - base : 256;  // records of this block
    if (i < m) {
        // slot order: the two children of the left child (or the left child itself if it is a
        // leaf), then those of the right child; every index below is a compile-time constant, so
        // the record stays in registers
        const float4* nd = nodes + i * 4;
        const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
        const int32_t c0 = __float_as_int(n3.x), c1 = __float_as_int(n3.y);
        struct Slot {
            float lx, ly, lz, hx, hy, hz;
            int32_t ref;
        };
        // the one or two entries a child contributes: its grandchildren, or itself if a leaf
        auto entries = [&](int32_t c, float4 b, float bz0, float bz1, Slot& e0, Slot& e1) {
            if (c >= 0) {
                const float4* cd = nodes + (uint64_t)c * 4;
                const float4 m0 = cd[0], m1 = cd[1], m2 = cd[2], m3 = cd[3];
                e0 = {m0.x, m0.z, m2.x, m0.y, m0.w, m2.y, __float_as_int(m3.x)};
                e1 = {m1.x, m1.z, m2.z, m1.y, m1.w, m2.w, __float_as_int(m3.y)};
            } else {
                e0 = {b.x, b.z, bz0, b.y, b.w, bz1, c};
                e1 = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, RT_EMPTY_ROOT};
            }
        };
        Slot a0, a1, b0, b1;
        entries(c0, n0, n2.x, n2.y, a0, a1);
        entries(c1, n1, n2.z, n2.w, b0, b1);
        const bool two = c0 >= 0;  // the left child contributes two entries
        const Slot e[4] = {a0, two ? a1 : b0, two ? b0 : b1, two ? b1 : a1};
        // (with a leaf left child, slot 3 takes a1: the empty entry)
        float4* o = stage + threadIdx.x * kPitch;
        o[0] = make_float4(e[0].lx, e[1].lx, e[2].lx, e[3].lx);
        o[1] = make_float4(e[0].hx, e[1].hx, e[2].hx, e[3].hx);
        o[2] = make_float4(e[0].ly, e[1].ly, e[2].ly, e[3].ly);
        o[3] = make_float4(e[0].hy, e[1].hy, e[2].hy, e[3].hy);
        o[4] = make_float4(e[0].lz, e[1].lz, e[2].lz, e[3].lz);
        o[5] = make_float4(e[0].hz, e[1].hz, e[2].hz, e[3].hz);
        o[6] = make_float4(__int_as_float(e[0].ref), __int_as_float(e[1].ref), __int_as_float(e[2].ref),
                           __int_as_float(e[3].ref));
        o[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    __syncthreads();
    float4* out = nodes4 + base * 8;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t q = j * 256 + threadIdx.x;  // float4 q of the block's output
        if (q < rows * 8) out[q] = stage[(q >> 3) * kPitch + (q & 7)];
    }
}

// Scene of one triangle: no internal node; the root is leaf slot 0.
__global__ void k_single_leaf_bounds(const float4* __restrict__ leafbox, float* __restrict__ bounds) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const float4 a = leafbox[0], b = leafbox[1];
    bounds[0] = a.x; bounds[1] = a.y; bounds[2] = a.z; bounds[3] = 0.0f;
    bounds[4] = b.x; bounds[5] = b.y; bounds[6] = b.z; bounds[7] = 0.0f;
}

__global__ void k_empty_bounds(float* __restrict__ bounds) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    bounds[0] = bounds[1] = bounds[2] = INFINITY;
    bounds[4] = bounds[5] = bounds[6] = -INFINITY;
    bounds[3] = bounds[7] = 0.0f;
}

__global__ void __launch_bounds__(256) k_copy_u32(const uint32_t* __restrict__ a,
                                                   uint32_t* __restrict__ b, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

// ---- export in the reference's record types (rendering/_raycaster.py:8-22) ----
// BVH_AABB (48 B) as 3 float4: (min.xyz, 0), (max.xyz, 0), (count, elements, 0, 0) as int bits.
RT_DEV void put_aabb(float4* out, uint32_t j, float lx, float ly, float lz, float hx, float hy,
                     float hz, int32_t count, int32_t elements) {
    out[3 * (uint64_t)j] = make_float4(lx, ly, lz, 0.0f);
    out[3 * (uint64_t)j + 1] = make_float4(hx, hy, hz, 0.0f);
    out[3 * (uint64_t)j + 2] = make_float4(__int_as_float(count), __int_as_float(elements), 0.0f, 0.0f);
}

// Breadth-first numbering of the traversal tree in one workgroup: node 0 is the root; the
// children of the internal nodes of one level are appended as pairs in the order of their parents
// (an exclusive scan over each chunk of 1024 nodes of the level), so the numbering is a fixed
// function of the tree. A node's record is written by its parent (box, and its traversal ref
// parked in `elements` with count -1) and completed when its own level is processed (count and
// elements). 2 L - 1 records for L leaves.
__global__ void __launch_bounds__(1024) k_export_nodes(const float4* __restrict__ nodes, int32_t root,
                                                       const float* __restrict__ bounds,
                                                       float4* __restrict__ out,
                                                       uint32_t* __restrict__ n_out) {
    __shared__ uint32_t scan[1024];
    const uint32_t tid = threadIdx.x;
    if (tid == 0)
        put_aabb(out, 0, bounds[0], bounds[1], bounds[2], bounds[4], bounds[5], bounds[6], -1, root);
    __syncthreads();
    uint32_t lo = 0, hi = 1, next = 1;
    while (lo < hi) {
        for (uint32_t base = lo; base < hi; base += 1024) {
            const uint32_t j = base + tid;
            const bool in = j < hi;
            const int32_t ref = in ? __float_as_int(out[3 * (uint64_t)j + 2].y) : 0;
            const uint32_t internal = (in && ref >= 0) ? 1u : 0u;
            scan[tid] = internal;
            __syncthreads();
            for (uint32_t off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
                const uint32_t v = tid >= off ? scan[tid - off] : 0u;
                __syncthreads();
                scan[tid] += v;
                __syncthreads();
            }
            const uint32_t before = scan[tid] - internal, total = scan[1023];
            if (internal) {
                const uint32_t cb = next + 2u * before;
                const float4* nd = nodes + (uint64_t)ref * 4;
                const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
                put_aabb(out, cb, n0.x, n0.z, n2.x, n0.y, n0.w, n2.y, -1, __float_as_int(n3.x));
                put_aabb(out, cb + 1, n1.x, n1.z, n2.z, n1.y, n1.w, n2.w, -1, __float_as_int(n3.y));
                out[3 * (uint64_t)j + 2] = make_float4(__int_as_float(0), __int_as_float((int32_t)cb), 0.0f, 0.0f);
            } else if (in) {
                out[3 * (uint64_t)j + 2] = make_float4(__int_as_float((int32_t)leaf_count(ref)),
                                                       __int_as_float((int32_t)leaf_first(ref)), 0.0f, 0.0f);
            }
            next += 2u * total;
            __syncthreads();
        }
        lo = hi;
        hi = next;
    }
    if (tid == 0 && n_out) *n_out = next;
}

__global__ void k_export_empty(float4* __restrict__ out, uint32_t* __restrict__ n_out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        put_aabb(out, 0, INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0, -1);
        if (n_out) *n_out = 1;
    }
}

// prim_slot[order[s]] = s: the triangle-record slot of each global primitive id
__global__ void __launch_bounds__(256) k_prim_slot(const uint32_t* __restrict__ order, uint32_t n,
                                                        uint32_t* __restrict__ inv) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) inv[order[s]] = s;
}

// BVH_Triangle records (64 B as 4 float4: v0, v1, v2, (index, mesh, 0, 0)) of mesh m, written at
// the slot of each triangle; vertex indices are clamped to n_vertices - 1 as the build does.
__global__ void __launch_bounds__(256) k_export_tris(rt_mesh_desc d, uint32_t m, uint32_t base,
                                                     const uint32_t* __restrict__ inv,
                                                     float4* __restrict__ out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.n_tris) return;
    const uint64_t s = inv[base + t];
    float4* o = out + 4 * s;
    for (int k = 0; k < 3; ++k) {
        uint32_t vi = d.indices ? (uint32_t)d.indices[3 * (uint64_t)t + k] : 3 * t + k;
        vi = vi < d.n_vertices ? vi : d.n_vertices - 1;
        const float* P = (const float*)((const char*)d.vertices + (uint64_t)vi * d.vertex_stride + d.pos_offset);
        o[k] = make_float4(P[0], P[1], P[2], 0.0f);
    }
    o[3] = make_float4(__int_as_float((int32_t)t), __int_as_float((int32_t)m), 0.0f, 0.0f);
}

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" int rt_scene_sizes_for(uint32_t n_tris, rt_scene_sizes* out) {
    if (!out) return set_error(RT_EINVAL, "rt_scene_sizes_for: out is NULL");
    if (n_tris > kMaxTris) return set_error(RT_ECAPACITY, "rt_scene_sizes_for: too many triangles");
    const uint64_t n = n_tris;
    out->nodes_bytes = 64 * (n > 1 ? n - 1 : 1);
    out->nodes4_bytes = 128 * (n > 1 ? n - 1 : 1);
    out->tris_bytes = 48 * (n ? n : 1);
    out->shade_bytes = 64 * (n ? n : 1);
    out->bounds_bytes = 32;
    out->workspace_bytes = carve(nullptr, n ? n : 1).bytes;
    out->prim_slot_bytes = 4 * (n ? n : 1);
    return RT_OK;
}

extern "C" int rt_scene_build(const rt_mesh_desc* meshes, uint32_t n_meshes, rt_scene* sc,
                              void* workspace, uint64_t ws_bytes, uint32_t* keys_out,
                              uint32_t* order_out, const rt_tuning* tuning, void* stream) {
    if (!sc) return set_error(RT_EINVAL, "rt_scene_build: scene is NULL");
    rt_tuning tu;
    if (int e = resolve_tuning(tuning, tu, "rt_scene_build")) return e;
    if (n_meshes && !meshes) return set_error(RT_EINVAL, "rt_scene_build: meshes is NULL");
    uint64_t n64 = 0;
    for (uint32_t m = 0; m < n_meshes; ++m) {
        const rt_mesh_desc& d = meshes[m];
        if (d.n_tris && !d.vertices) return set_error(RT_EINVAL, "rt_scene_build: mesh vertices NULL");
        if (d.n_tris && d.n_vertices == 0) return set_error(RT_EINVAL, "rt_scene_build: mesh has no vertices");
        if (!d.indices && (uint64_t)d.n_tris * 3 > d.n_vertices)
            return set_error(RT_EINVAL, "rt_scene_build: non-indexed mesh needs 3 vertices per triangle");
        if (d.vertex_stride < 12 || (d.vertex_stride & 3) || (d.pos_offset & 3) ||
            d.pos_offset + 12 > d.vertex_stride)
            return set_error(RT_EINVAL, "rt_scene_build: bad vertex stride/position offset");
        if (d.nrm_offset != RT_ABSENT && ((d.nrm_offset & 3) || d.nrm_offset + 12 > d.vertex_stride))
            return set_error(RT_EINVAL, "rt_scene_build: bad normal offset");
        if (d.uv_offset != RT_ABSENT && ((d.uv_offset & 3) || d.uv_offset + 8 > d.vertex_stride))
            return set_error(RT_EINVAL, "rt_scene_build: bad uv offset");
        n64 += d.n_tris;
    }
    if (n64 > kMaxTris) return set_error(RT_ECAPACITY, "rt_scene_build: too many triangles");
    const uint32_t n = (uint32_t)n64;
    if (sc->n_tris != n || sc->n_meshes != n_meshes)
        return set_error(RT_EINVAL, "rt_scene_build: scene n_tris/n_meshes do not match the meshes");
    if (!sc->nodes || !sc->nodes4 || !sc->tris || !sc->shade || !sc->mesh_offsets || !sc->bounds ||
        !sc->prim_slot)
        return set_error(RT_EINVAL, "rt_scene_build: scene arrays must be allocated");
    if (((uintptr_t)sc->nodes | (uintptr_t)sc->nodes4 | (uintptr_t)sc->tris | (uintptr_t)sc->shade) & 15)
        return set_error(RT_EINVAL, "rt_scene_build: scene arrays must be 16-byte aligned");
    hipStream_t s = (hipStream_t)stream;

    if (n == 0) {
        sc->root = RT_EMPTY_ROOT;
        hipLaunchKernelGGL(k_empty_bounds, dim3(1), dim3(1), 0, s, sc->bounds);
        hipLaunchKernelGGL(k_mesh_offsets_tail, dim3(1), dim3(1), 0, s, sc->mesh_offsets, 0u, 0u);
        for (uint32_t m = 0; m < n_meshes; ++m)
            hipLaunchKernelGGL(k_mesh_offsets_tail, dim3(1), dim3(1), 0, s, sc->mesh_offsets, m, 0u);
        hipLaunchKernelGGL(k_mesh_offsets_tail, dim3(1), dim3(1), 0, s, sc->mesh_offsets, n_meshes, 0u);
        return check_launch("rt_scene_build(empty)");
    }
    const Workspace need = carve(nullptr, n);
    if (!workspace || ws_bytes < need.bytes)
        return set_error(RT_EINVAL, "rt_scene_build: workspace too small (need " +
                                        std::to_string(need.bytes) + " bytes)");
    if ((uintptr_t)workspace & 255)
        return set_error(RT_EINVAL, "rt_scene_build: workspace must be 256-byte aligned");
    Workspace w = carve((char*)workspace, n);

    // 1. gather (one launch per mesh; the descriptor travels as a kernel argument)
    uint32_t base = 0;
    for (uint32_t m = 0; m < n_meshes; ++m) {
        const rt_mesh_desc& d = meshes[m];
        hipLaunchKernelGGL(k_gather, dim3(d.n_tris ? blocks_for(d.n_tris, 256) : 1), dim3(256), 0, s,
                           d, m, base, w.cent, w.leafbox, w.tri_u, w.shade_u, sc->mesh_offsets);
        base += d.n_tris;
    }
    hipLaunchKernelGGL(k_mesh_offsets_tail, dim3(1), dim3(1), 0, s, sc->mesh_offsets, n_meshes, n);
    // 2. centroid bounds
    const int nparts = (int)std::min<uint32_t>(kReduceBlocks, blocks_for(n, 256));
    hipLaunchKernelGGL(k_bounds_partial, dim3(nparts), dim3(256), 0, s, w.cent, n, w.partial);
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(256), 0, s, w.partial, nparts, w.cbounds);
    // 3. Morton codes
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(n, 256)), dim3(256), 0, s, w.cent, n, w.cbounds,
                       w.keys0, w.vals0);
    // 4. radix sort, 4 passes over the 30-bit keys
    uint32_t *ki = w.keys0, *vi = w.vals0;
    sort_pairs(ki, vi, w.keys1, w.vals1, n, 30, w.hist, w.totals, s);
    // after 4 passes the sorted arrays are back in keys0/vals0
    float4* nodes = (float4*)sc->nodes;
    bool ploc = n > 1 && tu.builder == RT_BUILD_PLOC;
    const uint32_t nchunk = blocks_for(n, kPlocChunk);
    if (ploc) {
        // 5. PLOC hierarchy (node boxes as it merges, scene bounds at the root), 6. leaf slots and
        // traversal refs, 7. leaf records in slot order, 8. 4-wide records
        hipLaunchKernelGGL(k_ploc_init, dim3(blocks_for(n, 256)), dim3(256), 0, s, vi, n, w.leafbox,
                           w.cid[0], w.cbox[0], w.pos_of, w.parent, w.state);
        const int iters = ploc_iterations(n);
        const uint32_t gnn = std::min<uint32_t>(kPlocGrid, blocks_for(n, kPlocTile));
        const uint32_t gch = std::min<uint32_t>(kPlocGrid, nchunk);
        for (int it = 0; it < iters; ++it) {
            hipLaunchKernelGGL(k_ploc_nn, dim3(gnn), dim3(kPlocTile), 0, s, it, w.state, w.cbox[0],
                               w.cbox[1], w.nn);
            hipLaunchKernelGGL(k_ploc_count, dim3(gch), dim3(256), 0, s, it, w.state, w.nn, w.part);
            hipLaunchKernelGGL(k_ploc_offsets, dim3(1), dim3(1024), 0, s, it, w.state, w.part);
            hipLaunchKernelGGL(k_ploc_compact, dim3(gch), dim3(256), 0, s, it, n, w.state, w.nn, w.part,
                               w.cid[0], w.cid[1], w.cbox[0], w.cbox[1], w.nleft, w.nright, w.nsize,
                               w.parent, nodes, sc->bounds);
        }
        hipLaunchKernelGGL(k_ploc_tail, dim3(1), dim3(1024), 0, s, iters & 1, n, w.state, w.nn, w.cid[0],
                           w.cid[1], w.cbox[0], w.cbox[1], w.nleft, w.nright, w.nsize, w.parent, nodes,
                           sc->bounds);
        // a tree deeper than the traversal stacks allow is replaced by the LBVH of the same keys;
        // this reads one word back, so a PLOC build waits for its stream here
        hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(1), 0, s, w.depth);
        hipLaunchKernelGGL(k_ploc_depth, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, w.parent, w.depth);
        uint32_t deep = 0;
        int e = check_hip(hipMemcpyAsync(&deep, w.depth, 4, hipMemcpyDeviceToHost, s), "rt_scene_build(PLOC depth)");
        if (!e) e = check_hip(hipStreamSynchronize(s), "rt_scene_build(PLOC depth)");
        if (e) return e;
        ploc = deep == 0;
    }
    if (ploc) {
        hipLaunchKernelGGL(k_ploc_leafroots, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, vi, w.parent,
                           w.nsize, w.nleft, w.nright, w.pos_of, w.scan);
        hipLaunchKernelGGL(k_scan_chunks, dim3(nchunk), dim3(256), 0, s, n, w.scan, w.part);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, nchunk, w.part);
        hipLaunchKernelGGL(k_scan_apply, dim3(nchunk), dim3(256), 0, s, n, w.scan, w.part);
        hipLaunchKernelGGL(k_ploc_slots, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, vi, w.parent,
                           w.nsize, w.nleft, w.nright, w.pos_of, w.scan, w.slot_order, w.slot_of, w.fslot);
        hipLaunchKernelGGL(k_ploc_refs, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, n, w.nleft,
                           w.nright, w.nsize, w.slot_of, w.fslot, nodes);
        hipLaunchKernelGGL(k_emit, dim3(blocks_for(4 * (uint64_t)n, 256)), dim3(256), 0, s, w.slot_order, n, w.tri_u,
                           w.shade_u, (float4*)sc->tris, (float4*)sc->shade);
        hipLaunchKernelGGL(k_wide, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, n,
                           (const float4*)nodes, (float4*)sc->nodes4);
        sc->root = n <= (uint32_t)kLeafMax ? leaf_ref(0, n) : 0;
        vi = w.slot_order;
    } else {
    // 5. hierarchy, 6. emit, 7. refit (LBVH, or a PLOC tree found too deep)
    if (n > 1) {
        hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(1), 0, s, w.top_count);
        hipLaunchKernelGGL(k_hierarchy, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, ki, n, nodes,
                           w.parent, w.rangej, w.split, w.top_count, w.top_list);
    }
    hipLaunchKernelGGL(k_emit, dim3(blocks_for(4 * (uint64_t)n, 256)), dim3(256), 0, s, vi, n, w.tri_u, w.shade_u,
                       (float4*)sc->tris, (float4*)sc->shade);
    if (n > 1) {
        const uint32_t nb = blocks_for(n, kRefitBlock);
        hipLaunchKernelGGL(k_refit_local, dim3(nb), dim3(256), 0, s, vi, n, w.leafbox, w.parent,
                           w.rangej, w.split, nodes, w.nodebox, sc->bounds, w.pre, w.suf, w.sparse);
        hipLaunchKernelGGL(k_block_sparse, dim3(1), dim3(1024), 0, s, w.sparse, nb, sparse_levels(nb));
        hipLaunchKernelGGL(k_refit_top, dim3(std::min<uint32_t>(blocks_for(n, 256), 1024)), dim3(256), 0, s,
                           vi, n, w.leafbox, w.rangej, w.split, nodes, w.nodebox, sc->bounds, w.top_list,
                           w.top_count, w.pre, w.suf, w.sparse, nb);
        hipLaunchKernelGGL(k_wide, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, n,
                           (const float4*)nodes, (float4*)sc->nodes4);
        sc->root = n <= (uint32_t)kLeafMax ? leaf_ref(0, n) : 0;
    } else {
        hipLaunchKernelGGL(k_single_leaf_bounds, dim3(1), dim3(1), 0, s, w.leafbox, sc->bounds);
        sc->root = ~0;
    }
    }
    hipLaunchKernelGGL(k_prim_slot, dim3(blocks_for(n, 256)), dim3(256), 0, s, vi, n,
                       (uint32_t*)sc->prim_slot);
    if (keys_out)
        hipLaunchKernelGGL(k_copy_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, ki, keys_out, n);
    if (order_out)
        hipLaunchKernelGGL(k_copy_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, vi, order_out, n);
    return check_launch("rt_scene_build");
}

extern "C" int rt_scene_export(const rt_scene* sc, const rt_mesh_desc* meshes, uint32_t n_meshes,
                               void* aabb_out, void* tris_out, uint32_t* n_nodes_out, void* stream) {
    if (!sc || !aabb_out) return set_error(RT_EINVAL, "rt_scene_export: NULL argument");
    if (((uintptr_t)aabb_out | (uintptr_t)tris_out) & 15)
        return set_error(RT_EINVAL, "rt_scene_export: outputs must be 16-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    float4* aabb = (float4*)aabb_out;
    const uint32_t n = sc->n_tris;
    if (sc->root == RT_EMPTY_ROOT || n == 0) {
        hipLaunchKernelGGL(k_export_empty, dim3(1), dim3(1), 0, s, aabb, n_nodes_out);
        return check_launch("rt_scene_export");
    }
    if (!tris_out || !sc->prim_slot || !sc->bounds || (sc->root >= 0 && !sc->nodes))
        return set_error(RT_EINVAL, "rt_scene_export: NULL argument");
    if (n_meshes != sc->n_meshes || (n_meshes && !meshes))
        return set_error(RT_EINVAL, "rt_scene_export: the meshes must be those the scene was built from");
    uint64_t total = 0;
    for (uint32_t m = 0; m < n_meshes; ++m) {
        if (meshes[m].n_tris && (!meshes[m].vertices || !meshes[m].n_vertices))
            return set_error(RT_EINVAL, "rt_scene_export: mesh vertices NULL");
        total += meshes[m].n_tris;
    }
    if (total != n) return set_error(RT_EINVAL, "rt_scene_export: mesh triangle counts do not match the scene");
    const uint32_t* inv = (const uint32_t*)sc->prim_slot;
    uint32_t base = 0;
    for (uint32_t m = 0; m < n_meshes; ++m) {
        if (meshes[m].n_tris)
            hipLaunchKernelGGL(k_export_tris, dim3(blocks_for(meshes[m].n_tris, 256)), dim3(256), 0, s,
                               meshes[m], m, base, inv, (float4*)tris_out);
        base += meshes[m].n_tris;
    }
    hipLaunchKernelGGL(k_export_nodes, dim3(1), dim3(1024), 0, s, (const float4*)sc->nodes, sc->root,
                       sc->bounds, aabb, n_nodes_out);
    return check_launch("rt_scene_export");
}
