// rt_build.hip — on-device LBVH construction for rendering/_raycaster.py:26-33
// (Raycaster.__init__ -> _build_ads, a stub in the reference that only counts triangles).
//
// Pipeline (one stream, no host synchronisation):
//   1. gather      per mesh: read MeshVertex P/N/C (rendering/_modeling.py:22-28) through the
//                  optional int32 index list, write unsorted tri/shade records, padded leaf box,
//                  AABB centroid.
//   2. bounds      two-stage min/max reduction of the centroids.
//   3. morton      30-bit Morton code of each centroid quantised to 1024^3 over those bounds.
//   4. sort        stable LSD radix sort of (code, primitive) in 4 passes of 8 bits; each pass is
//                  per-tile histogram -> per-digit row scan -> scatter; the scatter block scans
//                  the 256 digit totals itself and ranks keys in LDS with wave64 ballot/mbcnt
//                  multi-split so equal digits keep input order.
//   5. hierarchy   Karras 2012 binary radix tree over the extended keys (code, slot).
//   6. emit        permute tri/shade records into slot order.
//   7. refit       bottom-up from every leaf; the second thread to reach a node unions its
//                  children's boxes and continues upward. Nodes whose leaf range lies inside one
//                  2048-leaf block are refit by that block's workgroup through LDS counters; the
//                  ~0.6% spanning blocks by a second pass with device-scope acq_rel counters
//                  (each device-scope release writes back L2, so it is kept off the bulk).
// Every result is an integer or an exact FP32 min/max/sub, so the output is byte-identical to the
// CPU restatement under the same inputs.
#include <algorithm>
#include <utility>

#include "rt_common.h"

namespace rt {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;                       // items per thread per tile
constexpr int kSortTile = kSortThreads * kSortItems;  // 2048 keys per block
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
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
    uint32_t* flags;   // n-1
    float4* nodebox;   // 2(n-1)
    int32_t* top_list; // n-1: internal nodes whose leaf range spans several refit blocks
    uint32_t* top_count;
    uint64_t bytes;
};

inline uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

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
    w.flags = (uint32_t*)take(4 * ni);
    w.nodebox = (float4*)take(32 * ni);
    w.top_list = (int32_t*)take(4 * ni);
    w.top_count = (uint32_t*)take(4);
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

__global__ void __launch_bounds__(256) k_gather(rt_mesh_desc m, uint32_t mesh_id,
                                                 uint32_t prim_base, float4* __restrict__ cent,
                                                 float4* __restrict__ leafbox,
                                                 float4* __restrict__ tri_u,
                                                 float4* __restrict__ shade_u,
                                                 int32_t* __restrict__ mesh_offsets) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) mesh_offsets[mesh_id] = (int32_t)prim_base;
    if (t >= m.n_tris) return;
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
    leafbox[2 * (uint64_t)p] = make_float4(lox - px, loy - py, loz - pz, 0.0f);
    leafbox[2 * (uint64_t)p + 1] = make_float4(hix + px, hiy + py, hiz + pz, 0.0f);

    const float e1x = p1.x - p0.x, e1y = p1.y - p0.y, e1z = p1.z - p0.z;
    const float e2x = p2.x - p0.x, e2y = p2.y - p0.y, e2z = p2.z - p0.z;
    float4* tr = tri_u + 3 * (uint64_t)p;
    tr[0] = make_float4(p0.x, p0.y, p0.z, __int_as_float((int32_t)p));
    tr[1] = make_float4(e1x, e1y, e1z, 0.0f);
    tr[2] = make_float4(e2x, e2y, e2z, 0.0f);

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
    float4* sh = shade_u + 4 * (uint64_t)p;
    sh[0] = make_float4(n0.x, n0.y, n0.z, n1.x);
    sh[1] = make_float4(n1.y, n1.z, n2.x, n2.y);
    sh[2] = make_float4(n2.z, uv0.x, uv0.y, uv1.x);
    sh[3] = make_float4(uv1.y, uv2.x, uv2.y, __int_as_float((int32_t)mesh_id));
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

// ---- radix sort ----

__global__ void __launch_bounds__(kSortThreads) k_sort_hist(const uint32_t* __restrict__ keys,
                                                            uint32_t n, int shift,
                                                            uint32_t* __restrict__ hist,
                                                            uint32_t nblk) {
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        const uint64_t i = base + (uint64_t)r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// Row scan: block d turns row d of the digit-major histogram (counts of digit d in tiles
// 0..nblk-1) into its exclusive prefix sums, in place, and writes the row total to totals[d].
__global__ void __launch_bounds__(kSortThreads) k_sort_rowscan(uint32_t* __restrict__ hist,
                                                               uint32_t nblk,
                                                               uint32_t* __restrict__ totals) {
    __shared__ uint32_t s[kSortThreads];
    uint32_t* row = hist + (uint64_t)blockIdx.x * nblk;
    uint32_t run = 0;
    for (uint32_t c0 = 0; c0 < nblk; c0 += kSortThreads) {
        const uint32_t i = c0 + threadIdx.x;
        const uint32_t v = i < nblk ? row[i] : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < kSortThreads; off <<= 1) {
            const uint32_t x = threadIdx.x >= (uint32_t)off ? s[threadIdx.x - off] : 0u;
            __syncthreads();
            s[threadIdx.x] += x;
            __syncthreads();
        }
        if (i < nblk) row[i] = run + s[threadIdx.x] - v;
        run += s[kSortThreads - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = run;
}

// Stable scatter: items of a tile are visited in rounds of 256 (item order = (round, thread)),
// each round ranked by wave64 multi-split (8 ballots give the lanes sharing this lane's digit).
__global__ void __launch_bounds__(kSortThreads) k_sort_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t n, int shift,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ totals, uint32_t nblk,
    uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout) {
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint32_t base[kRadix];
    __shared__ uint32_t run[kRadix];
    __shared__ uint32_t wcount[kWaves][kRadix];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // exclusive scan of the digit totals gives each digit's first output position
    run[tid] = totals[tid];
    __syncthreads();
    for (int off = 1; off < kRadix; off <<= 1) {
        const uint32_t x = tid >= off ? run[tid - off] : 0u;
        __syncthreads();
        run[tid] += x;
        __syncthreads();
    }
    base[tid] = run[tid] - totals[tid] + hist[(uint64_t)tid * nblk + blockIdx.x];
    __syncthreads();
    run[tid] = 0;
    for (int w = 0; w < kWaves; ++w) wcount[w][tid] = 0;
    __syncthreads();
    const uint64_t tile = (uint64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        const uint64_t i = tile + (uint64_t)r * kSortThreads + tid;
        const bool valid = i < n;
        const uint32_t key = valid ? kin[i] : 0u;
        const uint32_t val = valid ? vin[i] : 0u;
        const uint32_t d = (key >> shift) & (kRadix - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi(
            (uint32_t)(peers >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)peers, 0u));
        const int leader = __ffsll((unsigned long long)peers) - 1;
        if (valid && lane == leader) wcount[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t off = base[d] + run[d] + rank;
            for (int w = 0; w < wave; ++w) off += wcount[w][d];
            if (off < n) {  // always true for a consistent histogram; keeps writes in bounds
                kout[off] = key;
                vout[off] = val;
            }
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kWaves; ++w) {
            add += wcount[w][tid];
            wcount[w][tid] = 0;
        }
        run[tid] += add;
        __syncthreads();
    }
}

// ---- hierarchy ----

// Refit blocks: bottom-up refit inside each range of kRefitBlock leaves runs in one workgroup
// (k_refit_local); the nodes whose range spans several blocks ("top" nodes) are refit afterwards.
constexpr int kRefitBlock = 2048;

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

__global__ void __launch_bounds__(256) k_emit(const uint32_t* __restrict__ order, uint32_t n,
                                               const float4* __restrict__ tri_u,
                                               const float4* __restrict__ shade_u,
                                               float4* __restrict__ tris,
                                               float4* __restrict__ shade) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint64_t p = order[s];
    tris[3 * (uint64_t)s] = tri_u[3 * p];
    tris[3 * (uint64_t)s + 1] = tri_u[3 * p + 1];
    tris[3 * (uint64_t)s + 2] = tri_u[3 * p + 2];
    shade[4 * (uint64_t)s] = shade_u[4 * p];
    shade[4 * (uint64_t)s + 1] = shade_u[4 * p + 1];
    shade[4 * (uint64_t)s + 2] = shade_u[4 * p + 2];
    shade[4 * (uint64_t)s + 3] = shade_u[4 * p + 3];
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

// Phase 1: block b walks up from each of its leaves through the local nodes (those whose leaf range
// lies inside the block; a node's id is an end of its range, so local nodes have ids in the block
// too). The block first copies what the walks read into LDS: the rangej / split / parent entries of
// its node ids and the boxes of its leaves (gathered through `order`). The first thread to reach a
// node stops; the second (LDS counter, workgroup acq_rel) unions the children's boxes from LDS and
// writes the node. Also clears the counters of this block's node ids for phase 2.
__global__ void __launch_bounds__(256) k_refit_local(const uint32_t* __restrict__ order, uint32_t n,
                                                      const float4* __restrict__ leafbox,
                                                      const int32_t* __restrict__ parent,
                                                      const int32_t* __restrict__ rangej,
                                                      const int32_t* __restrict__ split,
                                                      uint32_t* __restrict__ flags,
                                                      float4* __restrict__ nodes,
                                                      float4* __restrict__ nodebox,
                                                      float* __restrict__ bounds) {
    __shared__ uint32_t lflag[kRefitBlock];
    __shared__ float lbox[kRefitBlock][6];   // boxes of the block's internal nodes
    __shared__ float lleaf[kRefitBlock][6];  // boxes of the block's leaves (slots)
    __shared__ int32_t lj[kRefitBlock], lsplit[kRefitBlock], lpar[kRefitBlock];
    const int64_t blo = (int64_t)blockIdx.x * kRefitBlock;
    const int64_t bhi = std::min<int64_t>(blo + kRefitBlock, n);  // exclusive
    const int64_t ni = (int64_t)n - 1;
    for (int k = threadIdx.x; k < kRefitBlock; k += blockDim.x) {
        lflag[k] = 0;
        if (blo + k < ni) {
            flags[blo + k] = 0;
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

// Phase 2, in one workgroup, over the list of top nodes written by k_hierarchy: their local
// children are complete (phase 1); a node with two local children is computed at once, one with a
// single local child counts that child's arrival on the node's counter. The second arrival at a
// node computes it and continues upward (every ancestor of a top node is a top node). All counters
// are touched by the waves of this one workgroup on one CU, so workgroup-scope atomics order the
// box stores; an agent-scope release would write the L2 back on every step of the chain, the L2s
// of the 8 XCDs not being coherent with each other.
__global__ void __launch_bounds__(1024) k_refit_top(const uint32_t* __restrict__ order, uint32_t n,
                                                      const float4* __restrict__ leafbox,
                                                      const int32_t* __restrict__ parent,
                                                      const int32_t* __restrict__ rangej,
                                                      const int32_t* __restrict__ split,
                                                      uint32_t* __restrict__ flags,
                                                      float4* __restrict__ nodes, float4* nodebox,
                                                      float* __restrict__ bounds,
                                                      const int32_t* __restrict__ top_list,
                                                      const uint32_t* __restrict__ top_count) {
    const uint32_t cnt = *top_count;
    for (uint32_t t = threadIdx.x; t < cnt; t += blockDim.x) {
        int32_t node = top_list[t];
        {
            int32_t l, r;
            binary_children(rangej, split, node, l, r);
            const int n_local = (int)is_local(rangej, l) + (int)is_local(rangej, r);
            if (n_local == 0) continue;
            if (n_local == 1) {
                const uint32_t prev = __hip_atomic_fetch_add(&flags[node], 1u, __ATOMIC_ACQ_REL,
                                                             __HIP_MEMORY_SCOPE_WORKGROUP);
                if (prev == 0) continue;
            }
        }
        for (;;) {
            int32_t l, r;
            binary_children(rangej, split, node, l, r);
            float l_lo[3], l_hi[3], r_lo[3], r_hi[3], u_lo[3], u_hi[3];
            node_box(leafbox, order, nodebox, l, l_lo, l_hi);
            node_box(leafbox, order, nodebox, r, r_lo, r_hi);
            store_node(nodes, nodebox, bounds, node, l_lo, l_hi, r_lo, r_hi, u_lo, u_hi);
            const int32_t p = parent[node];
            if (p < 0) break;
            const uint32_t prev = __hip_atomic_fetch_add(&flags[p], 1u, __ATOMIC_ACQ_REL,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP);
            if (prev == 0) break;
            node = p;
        }
    }
}

__global__ void k_zero_word(uint32_t* w) { *w = 0u; }

// ---- 4-wide nodes ----
// nodes4[i] (8 x float4, 128 B) lists the grandchildren of binary node i reached through its
// internal children, and its leaf children directly, in the order (left..., right...):
//   [0]=lo.x[4] [1]=hi.x[4] [2]=lo.y[4] [3]=hi.y[4] [4]=lo.z[4] [5]=hi.z[4] [6]=refs[4] [7]=0
// Unused slots hold ref RT_EMPTY_ROOT and an empty box (lo=+inf, hi=-inf). Every binary node gets
// a record (traversal from the root only reaches every other level), so refs stay binary ids.
__global__ void __launch_bounds__(256) k_wide(uint32_t n, const float4* __restrict__ nodes,
                                               float4* __restrict__ nodes4) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n - 1) return;
    float lo[3][4], hi[3][4];
    int32_t ref[4];
    int k = 0;
    const float4* nd = nodes + (uint64_t)i * 4;
    const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
    const int32_t c[2] = {__float_as_int(n3.x), __float_as_int(n3.y)};
    const float cb[2][6] = {{n0.x, n0.z, n2.x, n0.y, n0.w, n2.y}, {n1.x, n1.z, n2.z, n1.y, n1.w, n2.w}};
    for (int e = 0; e < 2; ++e) {
        if (c[e] >= 0) {
            const float4* cd = nodes + (uint64_t)c[e] * 4;
            const float4 m0 = cd[0], m1 = cd[1], m2 = cd[2], m3 = cd[3];
            lo[0][k] = m0.x; hi[0][k] = m0.y; lo[1][k] = m0.z; hi[1][k] = m0.w;
            lo[2][k] = m2.x; hi[2][k] = m2.y; ref[k] = __float_as_int(m3.x); ++k;
            lo[0][k] = m1.x; hi[0][k] = m1.y; lo[1][k] = m1.z; hi[1][k] = m1.w;
            lo[2][k] = m2.z; hi[2][k] = m2.w; ref[k] = __float_as_int(m3.y); ++k;
        } else {
            lo[0][k] = cb[e][0]; lo[1][k] = cb[e][1]; lo[2][k] = cb[e][2];
            hi[0][k] = cb[e][3]; hi[1][k] = cb[e][4]; hi[2][k] = cb[e][5];
            ref[k] = c[e]; ++k;
        }
    }
    for (; k < 4; ++k) {
        for (int a = 0; a < 3; ++a) { lo[a][k] = INFINITY; hi[a][k] = -INFINITY; }
        ref[k] = RT_EMPTY_ROOT;
    }
    float4* o = nodes4 + (uint64_t)i * 8;
    for (int a = 0; a < 3; ++a) {
        o[2 * a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
        o[2 * a + 1] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
    }
    o[6] = make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), __int_as_float(ref[2]),
                       __int_as_float(ref[3]));
    o[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
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
    return RT_OK;
}

extern "C" int rt_scene_build(const rt_mesh_desc* meshes, uint32_t n_meshes, rt_scene* sc,
                              void* workspace, uint64_t ws_bytes, uint32_t* keys_out,
                              uint32_t* order_out, void* stream) {
    if (!sc) return set_error(RT_EINVAL, "rt_scene_build: scene is NULL");
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
    if (!sc->nodes || !sc->nodes4 || !sc->tris || !sc->shade || !sc->mesh_offsets || !sc->bounds)
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
    const uint32_t nblk = blocks_for(n, kSortTile);
    uint32_t *ki = w.keys0, *vi = w.vals0, *ko = w.keys1, *vo = w.vals1;
    for (int shift = 0; shift < 30; shift += kRadixBits) {
        hipLaunchKernelGGL(k_sort_hist, dim3(nblk), dim3(kSortThreads), 0, s, ki, n, shift, w.hist, nblk);
        hipLaunchKernelGGL(k_sort_rowscan, dim3(kRadix), dim3(kSortThreads), 0, s, w.hist, nblk,
                           w.totals);
        hipLaunchKernelGGL(k_sort_scatter, dim3(nblk), dim3(kSortThreads), 0, s, ki, vi, n, shift,
                           w.hist, w.totals, nblk, ko, vo);
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    // after 4 passes the sorted arrays are back in keys0/vals0
    // 5. hierarchy, 6. emit, 7. refit
    float4* nodes = (float4*)sc->nodes;
    if (n > 1) {
        hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(1), 0, s, w.top_count);
        hipLaunchKernelGGL(k_hierarchy, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, ki, n, nodes,
                           w.parent, w.rangej, w.split, w.top_count, w.top_list);
    }
    hipLaunchKernelGGL(k_emit, dim3(blocks_for(n, 256)), dim3(256), 0, s, vi, n, w.tri_u, w.shade_u,
                       (float4*)sc->tris, (float4*)sc->shade);
    if (n > 1) {
        hipLaunchKernelGGL(k_refit_local, dim3(blocks_for(n, kRefitBlock)), dim3(256), 0, s, vi, n,
                           w.leafbox, w.parent, w.rangej, w.split, w.flags, nodes, w.nodebox,
                           sc->bounds);
        hipLaunchKernelGGL(k_refit_top, dim3(1), dim3(1024), 0, s, vi, n, w.leafbox, w.parent,
                           w.rangej, w.split, w.flags, nodes, w.nodebox, sc->bounds, w.top_list,
                           w.top_count);
        hipLaunchKernelGGL(k_wide, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, n,
                           (const float4*)nodes, (float4*)sc->nodes4);
        sc->root = n <= (uint32_t)kLeafMax ? leaf_ref(0, n) : 0;
    } else {
        hipLaunchKernelGGL(k_single_leaf_bounds, dim3(1), dim3(1), 0, s, w.leafbox, sc->bounds);
        sc->root = ~0;
    }
    if (keys_out)
        hipLaunchKernelGGL(k_copy_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, ki, keys_out, n);
    if (order_out)
        hipLaunchKernelGGL(k_copy_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, vi, order_out, n);
    return check_launch("rt_scene_build");
}
