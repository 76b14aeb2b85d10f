// rt_trace.hip — closest-hit / any-hit queries and the render path.
//
// Queries (rendering/_raycaster.py:35-36, Raycaster.ray_cast — a stub in the reference):
//   k_query (closest or any hit): each lane walks its own ray with its own LDS stack (rays may be
//   incoherent) and takes the next ray of its wave's chunk when its walk ends; rays and hits are
//   read/written through (pointer, byte stride) fields, so SoA (stride 4) and the AoS Ray/Hit
//   records are both accepted.
// Camera rays:
//   each wave64 owns an 8x8 pixel tile (a 256-thread block = 16x16 pixels); rays are generated in
//   registers and traced with the wave-coherent traversal (scalar node fetches, one VGPR stack).
//   k_trace_camera writes per-pixel hit records.
//   k_render fuses the whole frame: primary closest hit, FP64 hit refinement, Lambert, one shadow
//   ray toward the light per lit hit (wave-coherent any-hit over the tile's hit points, skipping
//   each hit's own triangle), RGBA8 texture fetch, RGB float32 store.
#include <algorithm>
#include <mutex>

#include "rt_trace_shared.h"

namespace rt {
namespace {

RT_DEV float ld_f(const float* base, uint64_t i, uint32_t stride) {
    return *(const float*)((const char*)base + i * stride);
}

template <typename T>
RT_DEV void st(T* base, uint64_t i, uint32_t stride, T v) {
    if (base) *(T*)((char*)base + i * stride) = v;
}

RT_DEV int32_t mesh_of_prim(const int32_t* __restrict__ offs, uint32_t n_meshes, int32_t prim) {
    // largest m with offs[m] <= prim (empty meshes share offsets: take the last such m)
    uint32_t lo = 0, hi = n_meshes;  // answer in [lo, hi)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (offs[mid] <= prim) lo = mid; else hi = mid;
    }
    return (int32_t)lo;
}

RT_DEV Ray load_ray(const rt_rays& rays, uint64_t i) {
    Ray r;
    r.ox = ld_f(rays.ox, i, rays.stride);
    r.oy = ld_f(rays.oy, i, rays.stride);
    r.oz = ld_f(rays.oz, i, rays.stride);
    r.dx = ld_f(rays.dx, i, rays.stride);
    r.dy = ld_f(rays.dy, i, rays.stride);
    r.dz = ld_f(rays.dz, i, rays.stride);
    r.tmin = rays.tmin ? ld_f(rays.tmin, i, rays.stride) : rays.tmin_default;
    r.tmax = rays.tmax ? ld_f(rays.tmax, i, rays.stride) : rays.tmax_default;
    return r;
}

RT_DEV void store_hit(const SceneArgs& sc, const rt_hits& hits, uint64_t i, const Hit& h) {
    const bool hit = h.prim >= 0;
    int32_t mesh = -1, index = -1;
    if (hit && (hits.mesh || hits.index)) {
        mesh = mesh_of_prim(sc.mesh_offsets, sc.n_meshes, h.prim);
        index = h.prim - sc.mesh_offsets[mesh];
    }
    st(hits.t, i, hits.stride, hit ? h.t : INFINITY);
    st(hits.u, i, hits.stride, hit ? h.u : 0.0f);
    st(hits.v, i, hits.stride, hit ? h.v : 0.0f);
    st(hits.prim, i, hits.stride, h.prim);
    st(hits.index, i, hits.stride, index);
    st(hits.mesh, i, hits.stride, mesh);
}

// A ray of a query buffer, or a placeholder for lanes past the end (the traversals need every
// lane of a wave; the placeholder never takes part).
RT_DEV Ray query_ray(const rt_rays& rays, uint32_t i, bool valid) {
    if (valid) return load_ray(rays, i);
    Ray r;
    r.ox = r.oy = r.oz = 0.0f;
    r.dx = 1.0f;
    r.dy = r.dz = 0.0f;
    r.tmin = 0.0f;
    r.tmax = INFINITY;
    return r;
}

// Ray queries (closest hit / any hit of a ray buffer): per-lane walks with refill (refill_walks)
// over a grid of as many waves as the device holds at once.
template <bool ANY>
RT_DEV void finish_query(const SceneArgs& sc, const rt_hits& hits, uint8_t* occluded, uint32_t i,
                         const Ray& r, Hit& h) {
    if (ANY) {
        occluded[i] = h.prim >= 0 ? 1 : 0;
    } else {
        if (h.prim >= 0) refine_hit(sc.tris, r, h);
        store_hit(sc, hits, i, h);
    }
}

template <bool ANY, int MODE>
__global__ void __launch_bounds__(256) k_query(SceneArgs sc, rt_rays rays, uint32_t n, rt_hits hits,
                                                uint8_t* occluded, int lane_stack, uint32_t n_waves) {
    extern __shared__ int32_t qstk[];  // lane_stack entries per thread, entry k at [k * 256]
    __shared__ float4 pk[4][kPacketLds];
    __shared__ uint32_t ovf[4][kRefillList];
    const uint32_t wave = threadIdx.x >> 6;
    refill_walks<ANY, MODE, ANY ? RT_PARK_QUERY_ANY : RT_PARK_QUERY>(
        sc, n, n_waves, refill_slot(blockIdx.x, wave, 4, gridDim.x), qstk + threadIdx.x, 256, lane_stack, ovf[wave], pk[wave],
        [&](uint32_t i, Ray& r) -> int32_t {
            r = load_ray(rays, i);
            return -1;
        },
        [&](bool fin, uint32_t i, const Ray& r, Hit& h) {
            if (fin) finish_query<ANY>(sc, hits, occluded, i, r, h);
        });
}

__global__ void __launch_bounds__(256) k_generate_rays(rt_camera cam, uint32_t W, uint32_t H,
                                                        float* ox, float* oy, float* oz,
                                                        float* dx, float* dy, float* dz) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)W * H) return;
    float x, y, z;
    camera_dir(cam, W, H, (uint32_t)(i % W), (uint32_t)(i / W), x, y, z);
    ox[i] = cam.eye[0];
    oy[i] = cam.eye[1];
    oz[i] = cam.eye[2];
    dx[i] = x;
    dy[i] = y;
    dz[i] = z;
}

constexpr int kTile = 16;  // block = 16x16 pixels, wave = 8x8

struct TilePixel {
    uint32_t col, row;
    bool valid;
};

// Block -> tile mapping. Workgroups are handed to the 8 XCDs round robin (block b runs on XCD
// b % 8) and each XCD has its own L2, so consecutive row-major tiles would spread every screen
// region over all 8 L2s. Instead the image is cut into super-tiles of S x S blocks, super-tile s
// belongs to XCD s % 8, and XCD x walks its super-tiles in order: the rays one L2 serves come from
// a few compact screen regions and share their upper tree and surface nodes. S = 0 keeps the
// row-major order.
struct TileMap {
    uint32_t tiles_x, tiles_y;  // blocks (16x16 pixels; 8x8 for the quad camera kernel) across / down
    uint32_t S, super_x, n_super;
};

TileMap make_tile_map(uint32_t W, uint32_t H, uint32_t S, uint32_t tile = kTile) {
    TileMap m;
    m.tiles_x = (W + tile - 1) / tile;
    m.tiles_y = (H + tile - 1) / tile;
    m.S = S;
    m.super_x = S ? (m.tiles_x + S - 1) / S : 0;
    m.n_super = S ? m.super_x * ((m.tiles_y + S - 1) / S) : 0;
    return m;
}

uint32_t tile_map_blocks(const TileMap& m) {
    if (!m.S) return m.tiles_x * m.tiles_y;
    return 8u * m.S * m.S * ((m.n_super + 7u) / 8u);
}

// Tile (tx, ty) of block b; false if the block maps outside the image (it then does nothing).
RT_DEV bool block_tile_of(const TileMap& m, uint32_t b, uint32_t& tx, uint32_t& ty) {
    if (!m.S) {
        tx = b % m.tiles_x;
        ty = b / m.tiles_x;
        return ty < m.tiles_y;
    }
    const uint32_t x = b & 7u, j = b >> 3, SS = m.S * m.S;
    const uint32_t s = (j / SS) * 8u + x, w = j % SS;
    if (s >= m.n_super) return false;
    tx = (s % m.super_x) * m.S + w % m.S;
    ty = (s / m.super_x) * m.S + w / m.S;
    return tx < m.tiles_x && ty < m.tiles_y;
}

RT_DEV bool block_tile(const TileMap& m, uint32_t& tx, uint32_t& ty) { return block_tile_of(m, blockIdx.x, tx, ty); }

RT_DEV TilePixel tile_pixel(uint32_t W, uint32_t H, uint32_t tx, uint32_t ty) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    TilePixel p;
    p.col = tx * kTile + (wave & 1) * 8 + (lane & 7);
    p.row = ty * kTile + (wave >> 1) * 8 + (lane >> 3);
    p.valid = p.col < W && p.row < H;
    return p;
}

RT_DEV Ray camera_ray(const rt_camera& cam, uint32_t W, uint32_t H, const TilePixel& p) {
    Ray r;
    r.ox = cam.eye[0];
    r.oy = cam.eye[1];
    r.oz = cam.eye[2];
    const uint32_t col = p.valid ? p.col : 0, row = p.valid ? p.row : 0;
    camera_dir(cam, W, H, col, row, r.dx, r.dy, r.dz);
    r.tmin = 0.0f;
    r.tmax = INFINITY;
    return r;
}

// Camera closest hit: ray generation fused into the traversal, one hit record per pixel.
template <int MODE>
__global__ void __launch_bounds__(256) k_trace_camera(SceneArgs sc, rt_camera cam, uint32_t W,
                                                       uint32_t H, TileMap tm, int lane_stack,
                                                       int budget, rt_hits hits) {
    __shared__ int32_t stk[uses_lds(MODE) ? kLaneStack * 256 : 1];
    __shared__ float4 pk[4][kPacketLds];
    uint32_t tx, ty;
    if (!block_tile(tm, tx, ty)) return;
    const TilePixel p = tile_pixel(W, H, tx, ty);
    const Ray r = camera_ray(cam, W, H, p);
    Hit h;
    trace_block<false, MODE>(sc, r, p.valid, -1, h, pk[threadIdx.x >> 6], stk + threadIdx.x, 256,
                             lane_stack, budget);
    if (h.prim >= 0) refine_hit(sc.tris, r, h);
    if (p.valid) store_hit(sc, hits, (uint64_t)p.row * W + p.col, h);
}

// Camera closest hit with quad-lane packets, one wave per workgroup: wave j of XCD x (workgroup
// 8j + x, as workgroups are handed to the 8 XCDs round robin) takes the 4x4 quarter w = j & 3 of
// the 8x8-pixel block (j >> 2) * 8 + x of the tile map, at ((w & 1) * 4, (w >> 1) * 4), and lanes
// 4r..4r+3 the pixel (r & 3, r >> 2) of the quarter. The four quarters of a block are consecutive waves of one XCD
// (they share its L2), and a wave that ends frees its slot for the next one at once (a CU starts a
// 4-wave workgroup only once it has 4 free slots: 3-4% slower frames).
__global__ void __launch_bounds__(64) k_trace_camera_quad(SceneArgs sc, rt_camera cam, uint32_t W,
                                                          uint32_t H, TileMap tm, rt_hits hits) {
    const uint32_t lane = threadIdx.x & 63, qr = lane >> 2;
    const uint32_t xcd = blockIdx.x & 7u, j = blockIdx.x >> 3, wave = j & 3u;
    uint32_t tx, ty;
    if (!block_tile_of(tm, (j >> 2) * 8u + xcd, tx, ty)) return;
    TilePixel p;
    p.col = tx * 8 + (wave & 1) * 4 + (qr & 3);
    p.row = ty * 8 + (wave >> 1) * 4 + (qr >> 2);
    p.valid = p.col < W && p.row < H;
    const Ray r = camera_ray(cam, W, H, p);
    Hit h;
    traverse_quad<false>(sc.nodes4, sc.tris, sc.root, r, p.valid, -1, h);
    if (p.valid && (lane & 3) == 0) {
        if (h.prim >= 0) refine_hit(sc.tris, r, h);
        store_hit(sc, hits, (uint64_t)p.row * W + p.col, h);
    }
}

RT_DEV int wrap_i(int i, int n) {
    const int r = i % n;
    return r < 0 ? r + n : r;
}

RT_DEV float4 texel(const uint8_t* __restrict__ tex, int W, int x, int y) {
    const uint32_t p = *as_global((const uint32_t*)(tex + ((uint64_t)y * W + x) * 4));
    return make_float4((float)(p & 0xFF), (float)((p >> 8) & 0xFF), (float)((p >> 16) & 0xFF),
                       (float)(p >> 24));
}

// RGBA8 fetch: OpenCL 1.2 linear filter with repeat addressing, or point sampling with the
// reference's wrap rule (rendering/_core.py:94-96).
RT_DEV void tex_sample(const rt_texture& tx, float u, float v, bool bilinear, float* rgb) {
    const int W = (int)tx.width, H = (int)tx.height;
    if (W <= 0 || H <= 0 || !tx.texels) return;  // no texels: albedo stays 1
    const float su = u - floorf(u), sv = v - floorf(v);
    if (!bilinear) {
        int i = (int)(su * (float)W), j = (int)(sv * (float)H);
        i = i > W - 1 ? W - 1 : i;
        j = j > H - 1 ? H - 1 : j;
        const float4 c = texel(tx.texels, W, i, j);
        rgb[0] = c.x * (1.0f / 255.0f);
        rgb[1] = c.y * (1.0f / 255.0f);
        rgb[2] = c.z * (1.0f / 255.0f);
        return;
    }
    const float x = su * (float)W - 0.5f, y = sv * (float)H - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    const int i0 = wrap_i((int)fx, W), j0 = wrap_i((int)fy, H);
    const int i1 = wrap_i((int)fx + 1, W), j1 = wrap_i((int)fy + 1, H);
    const float4 t00 = texel(tx.texels, W, i0, j0), t10 = texel(tx.texels, W, i1, j0);
    const float4 t01 = texel(tx.texels, W, i0, j1), t11 = texel(tx.texels, W, i1, j1);
    const float c00[3] = {t00.x, t00.y, t00.z}, c10[3] = {t10.x, t10.y, t10.z};
    const float c01[3] = {t01.x, t01.y, t01.z}, c11[3] = {t11.x, t11.y, t11.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float top = fmaf(a, c10[c] - c00[c], c00[c]);
        const float bot = fmaf(a, c11[c] - c01[c], c01[c]);
        rgb[c] = fmaf(b, bot - top, top) * (1.0f / 255.0f);
    }
}

// Shading terms of a refined hit h: Lambert factor max(0, N.L) with the normal interpolated from the
// vertex normals and normalised, the interpolated texture coordinates and the mesh id.
RT_DEV void shade_terms(const SceneArgs& sc, const rt_render_params& p, const Hit& h, float& ndl, float& tu,
                        float& tv, int32_t& mesh) {
    const float4* s = sc.shade + (uint64_t)h.slot * 4;
    const float4 s0 = s[0], s1 = s[1], s2 = s[2], s3 = s[3];
    const float w0 = 1.0f - h.u - h.v;
    const float nx = fmaf(w0, s0.x, fmaf(h.u, s0.w, h.v * s1.z));
    const float ny = fmaf(w0, s0.y, fmaf(h.u, s1.x, h.v * s1.w));
    const float nz = fmaf(w0, s0.z, fmaf(h.u, s1.y, h.v * s2.x));
    const float l2 = fmaf(nx, nx, fmaf(ny, ny, nz * nz));
    const float il = l2 > 0.0f ? 1.0f / sqrtf(l2) : 0.0f;
    ndl = fmaf(nx * il, p.light[0], fmaf(ny * il, p.light[1], (nz * il) * p.light[2]));
    ndl = fmaxf(ndl, 0.0f);
    tu = fmaf(w0, s2.y, fmaf(h.u, s2.w, h.v * s3.y));
    tv = fmaf(w0, s2.z, fmaf(h.u, s3.x, h.v * s3.z));
    mesh = __float_as_int(s3.w);
}

// The shadow ray of hit h of primary ray r: from the hit point toward the light.
RT_DEV Ray shadow_ray(const rt_render_params& p, const Ray& r, const Hit& h) {
    Ray sr;
    sr.ox = fmaf(h.t, r.dx, r.ox);
    sr.oy = fmaf(h.t, r.dy, r.oy);
    sr.oz = fmaf(h.t, r.dz, r.oz);
    sr.dx = p.light[0];
    sr.dy = p.light[1];
    sr.dz = p.light[2];
    sr.tmin = p.shadow_tmin;
    sr.tmax = INFINITY;
    return sr;
}

// Albedo of a hit on mesh `mesh` at (tu, tv): its texture's sample, or 1 without one.
RT_DEV void albedo(const rt_render_params& p, int32_t mesh, float tu, float tv, float* alb) {
    if (!p.mesh_texture_dev) return;
    const int32_t ti = as_global(p.mesh_texture_dev)[mesh];
    if (ti >= 0 && (uint32_t)ti < p.n_textures)
        tex_sample(as_global(p.textures_dev)[ti], tu, tv, (p.flags & RT_RENDER_BILINEAR) != 0, alb);
}

// Shading of one pixel per lane: primary closest hit, FP64 refinement, Lambert with interpolated
// normals, one shadow ray toward the light per lit hit (skipping the hit's own triangle), RGBA8
// albedo, RGB store; per-wave counts added into stats (global, or a block's LDS counters). PM / SM:
// traversal (RT_TRAVERSAL_*) of the primary / shadow rays; `stk` = this thread's lane stack (entries
// `stk_stride` words apart). Called by all 64 lanes of a wave.
template <int PM, int SM>
RT_DEV void render_pixel(const SceneArgs& sc, const rt_camera& cam, const rt_render_params& p,
                         const TilePixel& px, float* __restrict__ rgb, float* __restrict__ out_t,
                         int32_t* __restrict__ out_prim, uint32_t* __restrict__ stats, float4* pk,
                         int32_t* stk, int stk_stride, int lane_stack, int budget, bool coherent) {
    // PM == RT_TRAVERSAL_QUAD: quad layout (the 4 lanes of a quad hold the same pixel), both rays by
    // the quad-lane traversal, outputs and counts from lane 0 of each quad
    constexpr bool QL = PM == RT_TRAVERSAL_QUAD;
    const uint32_t W = p.width, H = p.height;
    const Ray r = camera_ray(cam, W, H, px);
    Hit h;
    if constexpr (QL) {
        traverse_quad<false>(sc.nodes4, sc.tris, sc.root, r, px.valid, -1, h);
    } else {
        trace_primary<PM>(sc, r, px.valid, h, pk, stk, stk_stride, lane_stack, budget, coherent);
    }
    const bool hit = px.valid && h.prim >= 0;
    if (hit) refine_hit(sc.tris, r, h);

    float ndl = 0.0f, tu = 0.0f, tv = 0.0f;
    int32_t mesh = 0;
    if (hit) shade_terms(sc, p, h, ndl, tu, tv, mesh);

    // shadow ray toward the light from every lit hit of the tile
    const bool want_shadow = hit && (p.flags & RT_RENDER_SHADOWS) && ndl > 0.0f;
    const Ray sr = shadow_ray(p, r, h);
    Hit sh;
    if constexpr (QL)
        traverse_quad<true>(sc.nodes4, sc.tris, sc.root, sr, want_shadow, hit ? h.slot : -1, sh);
    else
        trace_block<true, SM>(sc, sr, want_shadow, hit ? h.slot : -1, sh, pk, stk, stk_stride, lane_stack,
                              budget);
    const bool occluded = want_shadow && sh.prim >= 0;

    float alb[3] = {1.0f, 1.0f, 1.0f};
    if (hit) albedo(p, mesh, tu, tv, alb);
    if (px.valid && (!QL || (threadIdx.x & 3) == 0)) {
        const uint64_t pix = (uint64_t)px.row * W + px.col;
        if (hit) {
            const float kk = fmaf(occluded ? 0.0f : 1.0f, ndl, p.ambient);
            rgb[3 * pix] = alb[0] * kk;
            rgb[3 * pix + 1] = alb[1] * kk;
            rgb[3 * pix + 2] = alb[2] * kk;
        } else {
            rgb[3 * pix] = p.background[0];
            rgb[3 * pix + 1] = p.background[1];
            rgb[3 * pix + 2] = p.background[2];
        }
        if (out_t) out_t[pix] = hit ? h.t : INFINITY;
        if (out_prim) out_prim[pix] = hit ? h.prim : -1;
    }
    if (stats) {
        const uint64_t one = QL ? kQuadBits : ~0ull;  // one lane per pixel
        const uint32_t n_hit = (uint32_t)__popcll(__ballot(hit) & one);
        const uint32_t n_sh = (uint32_t)__popcll(__ballot(want_shadow) & one);
        const uint32_t n_oc = (uint32_t)__popcll(__ballot(occluded) & one);
        if ((threadIdx.x & 63) == 0) {
            if (n_hit) atomicAdd(&stats[0], n_hit);
            if (n_sh) atomicAdd(&stats[1], n_sh);
            if (n_oc) atomicAdd(&stats[2], n_oc);
        }
    }
}

// The render counts of a block: its waves add theirs into `bstat` (LDS), and one thread per
// counter adds the block's total to `stats` (global), so a 1080p frame takes ~100K global atomics
// on the same three words instead of one per wave (~390K with 16-pixel waves). Called by all
// threads of the block with `stats` uniform.
RT_DEV void block_stats_begin(const uint32_t* stats, uint32_t* bstat) {
    if (!stats) return;
    if (threadIdx.x < 3) bstat[threadIdx.x] = 0;
    __syncthreads();
}

RT_DEV void block_stats_end(uint32_t* stats, uint32_t* bstat) {
    if (!stats) return;
    __syncthreads();
    if (threadIdx.x < 3 && bstat[threadIdx.x]) atomicAdd(&stats[threadIdx.x], bstat[threadIdx.x]);
}

// Fused render of one frame (see the file comment), one 16x16-pixel block per workgroup.
template <int PM, int SM>
__global__ void __launch_bounds__(256) k_render(SceneArgs sc, rt_camera cam, rt_render_params p,
                                                 TileMap tm, int lane_stack, int budget,
                                                 float* __restrict__ rgb, float* __restrict__ out_t,
                                                 int32_t* __restrict__ out_prim,
                                                 uint32_t* __restrict__ stats) {
    __shared__ int32_t stk[(uses_lds(PM) || uses_lds(SM)) ? kLaneStack * 256 : 1];
    __shared__ float4 pk[4][kPacketLds];
    __shared__ uint32_t bstat[3];
    uint32_t tx, ty;
    if (!block_tile(tm, tx, ty)) return;
    const TilePixel px = tile_pixel(p.width, p.height, tx, ty);
    block_stats_begin(stats, bstat);
    render_pixel<PM, SM>(sc, cam, p, px, rgb, out_t, out_prim, stats ? bstat : nullptr, pk[threadIdx.x >> 6],
                         stk + threadIdx.x, 256, lane_stack, budget, true);
    block_stats_end(stats, bstat);
}

// Fused render with quad-lane packets for the primary and the shadow rays: one 8x8-pixel tile per
// block, wave w = 4x4 quarter (w & 1, w >> 1), lanes 4q..4q+3 = pixel (q & 3, q >> 2).
// W1: one wave per workgroup (the camera kernel's mapping: wave j of XCD x renders quarter j & 3 of
// 8x8 block (j >> 2) * 8 + x, row-major); its counts, if any, go to partial slot blockIdx % 256 of
// `stats` (kCountSlots slots kCountStride words apart, summed by k_sum_counts), so that no counter
// takes more than one atomic in 256 of the frame's waves. Else 4-wave workgroups whose counts are
// summed in LDS and added to `stats` itself once per workgroup (for a render without a workspace).
constexpr uint32_t kCountSlots = 256, kCountStride = 32;  // 128 B per slot

template <bool W1>
__global__ void __launch_bounds__(W1 ? 64 : 256) k_render_quad(SceneArgs sc, rt_camera cam, rt_render_params p,
                                                      uint32_t tiles_x8, float* __restrict__ rgb,
                                                      float* __restrict__ out_t,
                                                      int32_t* __restrict__ out_prim,
                                                      uint32_t* __restrict__ stats) {
    const uint32_t q = (threadIdx.x & 63) >> 2;
    uint32_t wave, blk;
    if (W1) {
        const uint32_t xcd = blockIdx.x & 7u, j = blockIdx.x >> 3;
        wave = j & 3u;
        blk = (j >> 2) * 8u + xcd;
        if (blk >= tiles_x8 * ((p.height + 7u) / 8u)) return;
    } else {
        wave = threadIdx.x >> 6;
        blk = blockIdx.x;
    }
    TilePixel px;
    px.col = (blk % tiles_x8) * 8 + (wave & 1) * 4 + (q & 3);
    px.row = (blk / tiles_x8) * 8 + (wave >> 1) * 4 + (q >> 2);
    px.valid = px.col < p.width && px.row < p.height;
    if (W1) {
        uint32_t* slot = stats ? stats + (blockIdx.x % kCountSlots) * kCountStride : nullptr;
        render_pixel<RT_TRAVERSAL_QUAD, RT_TRAVERSAL_QUAD>(sc, cam, p, px, rgb, out_t, out_prim, slot, nullptr,
                                                           nullptr, 0, 0, 0, true);
    } else {
        __shared__ uint32_t bstat[3];
        block_stats_begin(stats, bstat);
        render_pixel<RT_TRAVERSAL_QUAD, RT_TRAVERSAL_QUAD>(sc, cam, p, px, rgb, out_t, out_prim,
                                                           stats ? bstat : nullptr, nullptr, nullptr, 0, 0, 0, true);
        block_stats_end(stats, bstat);
    }
}

// ---- batched render ----
// Workspace: [0] uint32 work counter, [1] status, then (n_jobs + 1) uint32 tile prefix at word 64.
constexpr uint32_t kBatchPrefixWord = 64;

// 8x8 tiles of a job, or 0 with a status bit if the job cannot be rendered: 1 = more than
// RT_BATCH_MAX_JOB_PIXELS pixels, 2 = a required pointer is NULL (rgb output, texture table with
// n_textures > 0, scene arrays of a non-empty scene).
RT_DEV uint32_t job_tiles(const rt_render_job& j, uint32_t unit, uint32_t& status) {
    const rt_render_params& p = j.params;
    const uint64_t px = (uint64_t)p.width * p.height;
    if (px == 0) return 0;
    if (px > RT_BATCH_MAX_JOB_PIXELS) {
        status |= 1u;
        return 0;
    }
    const rt_scene& s = j.scene;
    if (!j.out.rgb || (p.n_textures && !p.textures_dev) ||
        (s.root != RT_EMPTY_ROOT && (!s.nodes4 || !s.tris || !s.shade))) {
        status |= 2u;
        return 0;
    }
    return ((p.width + unit - 1u) / unit) * ((p.height + unit - 1u) / unit);
}

// One block of 1024 threads: per-job tile counts, exclusive prefix (n_jobs <= 4096, at most 4 jobs
// per thread), work counter and status cleared, every job's stats zeroed.
__global__ void __launch_bounds__(1024) k_batch_prepare(const rt_render_job* __restrict__ jobs,
                                                         uint32_t n_jobs, uint32_t unit,
                                                         uint32_t* __restrict__ ws) {
    __shared__ uint32_t part[1024];
    __shared__ uint32_t status;
    const uint32_t t = threadIdx.x;
    if (t == 0) status = 0;
    __syncthreads();
    uint32_t cnt[4], sum = 0, bad = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t j = t * 4 + k;
        cnt[k] = 0;
        if (j < n_jobs) {
            cnt[k] = job_tiles(jobs[j], unit, bad);
            uint32_t* st = as_global(jobs[j].out.stats);
            if (st) {
                st[0] = 0;
                st[1] = 0;
                st[2] = 0;
                st[3] = 0;
            }
        }
        sum += cnt[k];
    }
    part[t] = sum;
    __syncthreads();
    // Hillis-Steele inclusive scan over the 1024 partial sums
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    uint32_t* pref = ws + kBatchPrefixWord;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t j = t * 4 + k;
        if (j < n_jobs) pref[j] = run;
        run += cnt[k];
    }
    if (t == 1023) pref[n_jobs] = part[1023];
    if (bad) atomicOr(&status, bad);
    __syncthreads();
    if (t == 0) {
        ws[0] = 0;
        ws[1] = status;
    }
}

// A job's scene arrays, parameters and outputs with their pointers known to be global (as_global).
RT_DEV SceneArgs job_scene(const rt_render_job& j) {
    SceneArgs sc;
    sc.nodes = as_global((const float4*)j.scene.nodes);
    sc.nodes4 = as_global((const float4*)j.scene.nodes4);
    sc.tris = as_global((const float4*)j.scene.tris);
    sc.shade = as_global((const float4*)j.scene.shade);
    sc.mesh_offsets = as_global(j.scene.mesh_offsets);
    sc.prim_slot = as_global(j.scene.prim_slot);
    sc.root = j.scene.root;
    sc.n_meshes = j.scene.n_meshes;
    return sc;
}

RT_DEV rt_render_params job_params(const rt_render_job& j) {
    rt_render_params p = j.params;
    p.textures_dev = as_global(p.textures_dev);
    p.mesh_texture_dev = as_global(p.mesh_texture_dev);
    return p;
}

RT_DEV rt_render_out job_out(const rt_render_job& j) {
    rt_render_out o = j.out;
    o.rgb = as_global(o.rgb);
    o.t = as_global(o.t);
    o.prim = as_global(o.prim);
    o.stats = as_global(o.stats);
    return o;
}

// Persistent batched render: one wave per block; each wave pops `grab` consecutive 8x8 tiles of the
// concatenated job list from the work counter until the list is exhausted (every wave reaches the
// exit: the counter only grows). Tile -> job by binary search of the prefix, copied into LDS.
template <int PM, int SM>
__global__ void __launch_bounds__(64) k_render_batch(const rt_render_job* __restrict__ jobs,
                                                      uint32_t n_jobs, uint32_t* __restrict__ ws,
                                                      uint32_t grab, int lane_stack, int budget) {
    extern __shared__ uint32_t smem[];
    __shared__ float4 pk[kPacketLds];
    __shared__ uint32_t bstat[3];  // the wave's counts for job `sjob` since its last flush
    uint32_t* pref = smem;
    int32_t* stk = (int32_t*)(smem + ((n_jobs + 1 + 3) & ~3u));
    const uint32_t lane = threadIdx.x;
    for (uint32_t j = lane; j <= n_jobs; j += 64) pref[j] = ws[kBatchPrefixWord + j];
    if (lane < 3) bstat[lane] = 0;
    __syncthreads();
    const uint32_t total = pref[n_jobs];
    uint32_t* sjob = nullptr;  // wave-uniform: the stats words the counts belong to
    auto flush = [&]() {  // one atomic per counter per job the wave worked on, not one per tile
        __syncthreads();
        if (sjob && lane < 3 && bstat[lane]) atomicAdd(&sjob[lane], bstat[lane]);
        if (lane < 3) bstat[lane] = 0;
        __syncthreads();
    };
    for (;;) {
        uint32_t u0 = 0;
        if (lane == 0) u0 = atomicAdd(&ws[0], grab);
        u0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)u0);
        if (u0 >= total) break;
        const uint32_t u1 = u0 + grab < total ? u0 + grab : total;
        for (uint32_t u = u0; u < u1; ++u) {
            // largest j with pref[j] <= u (jobs with no tiles share a prefix value: take the last)
            uint32_t lo = 0, hi = n_jobs;  // answer in [lo, hi)
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (pref[mid] <= u) lo = mid; else hi = mid;
            }
            const uint32_t j = (uint32_t)__builtin_amdgcn_readfirstlane((int)lo);
            const rt_render_job& job = jobs[j];
            const rt_render_params p = job.params;
            constexpr uint32_t unit = PM == RT_TRAVERSAL_QUAD ? 4u : 8u;
            const uint32_t tiles_x = (p.width + unit - 1u) / unit, local = u - pref[j];
            // 8x8 tile: lane l -> (l & 7, l >> 3); quad layout 4x4 tile: lanes 4q..4q+3 -> (q & 3, q >> 2)
            const uint32_t dx = PM == RT_TRAVERSAL_QUAD ? (lane >> 2) & 3u : lane & 7u;
            const uint32_t dy = PM == RT_TRAVERSAL_QUAD ? lane >> 4 : lane >> 3;
            TilePixel px;
            px.col = (local % tiles_x) * unit + dx;
            px.row = (local / tiles_x) * unit + dy;
            px.valid = px.col < p.width && px.row < p.height;
            const SceneArgs sc = job_scene(job);
            const rt_render_out o = job_out(job);
            if (o.stats != sjob) {
                flush();
                sjob = o.stats;
            }
            render_pixel<PM, SM>(sc, job.cam, job_params(job), px, o.rgb, o.t, o.prim, sjob ? bstat : nullptr, pk,
                                 stk + lane, 64, lane_stack, budget, coherent_tiles(job.cam, p.width, p.height));
        }
    }
    flush();
}

// Batched render with per-lane refill (rt_batch_config primary = shadow = RT_TRAVERSAL_LANE). The
// wave takes `grab` tiles at a time from the work counter, as k_render_batch does, but its lanes
// are not tied to one tile: each lane carries one pixel through its primary walk, its shading and
// its shadow walk (both walks resumable, in rounds of kRefillSteps iterations), and whenever
// kRefillIdle or more lanes are idle they take the next pixels of the wave's tiles (one ballot).
// A wave holds pixels of one job at a time (it drains before it moves to another job's tiles), so
// the job's scene and parameters stay wave-uniform. A walk that would overflow its LDS stack is
// traced again at once by the whole wave with the packet traversal. Every pixel gets exactly the
// values render_pixel<LANE, LANE> gives it (same walks, refinement, shading and store).
#ifndef RT_REFILL_PASSES
#define RT_REFILL_PASSES 8
#endif
constexpr int kRefillPasses = RT_REFILL_PASSES;  // refills per round while culled rays free lanes

__global__ void __launch_bounds__(64) k_render_batch_refill(const rt_render_job* __restrict__ jobs,
                                                             uint32_t n_jobs, uint32_t* __restrict__ ws,
                                                             uint32_t grab, int lane_stack) {
    extern __shared__ uint32_t smem[];
    __shared__ float4 pk[kPacketLds];
    uint32_t* pref = smem;
    int32_t* stk = (int32_t*)(smem + ((n_jobs + 1 + 3) & ~3u)) + threadIdx.x;
    const uint32_t lane = threadIdx.x;
    for (uint32_t j = lane; j <= n_jobs; j += 64) pref[j] = ws[kBatchPrefixWord + j];
    __syncthreads();
    const uint32_t total = pref[n_jobs];

    // wave-uniform: the tile being handed out, its next pixel, the end of the grabbed tiles, and
    // the job the lanes work on (with its scene, camera, parameters and counts)
    uint32_t tile = 0, tile_px = 64, pool_end = 0, job = 0xFFFFFFFFu, tile_job = 0;
    bool exhausted = false;
    SceneArgs sc{};
    rt_camera cam{};
    rt_render_params p{};
    rt_render_out out{};
    uint32_t tiles_x = 1, job_first = 0, n_hit = 0, n_sh = 0, n_oc = 0;
    bool cull = false;  // the job's scene box culls new primary rays
    float blo0 = 0.0f, blo1 = 0.0f, blo2 = 0.0f, bhi0 = 0.0f, bhi1 = 0.0f, bhi2 = 0.0f;
    auto flush = [&]() {
        if (job != 0xFFFFFFFFu && out.stats && lane == 0) {
            if (n_hit) atomicAdd(&out.stats[0], n_hit);
            if (n_sh) atomicAdd(&out.stats[1], n_sh);
            if (n_oc) atomicAdd(&out.stats[2], n_oc);
        }
        n_hit = n_sh = n_oc = 0;
    };

    // per lane: the pixel, the phase (0: primary walk, 1: shadow walk), the walk and the shading
    bool busy = false;
    uint32_t phase = 0, col = 0, row = 0;
    Ray r;
    r.ox = r.oy = r.oz = 0.0f;
    r.dx = 1.0f;
    r.dy = r.dz = 0.0f;
    r.tmin = 0.0f;
    r.tmax = INFINITY;
    Hit h;
    h.t = INFINITY;
    h.u = h.v = 0.0f;
    h.prim = h.slot = -1;
    LaneWalk w;
    walk_start(r, RT_EMPTY_ROOT, w);
    int32_t skip = -1;
    float ndl = 0.0f, alb0 = 1.0f, alb1 = 1.0f, alb2 = 1.0f;

    for (;;) {
        // ---- refill from the current tile (moving on to the next tile / pool / job as needed). A new
        // primary ray that misses the scene's box (`cull`) is finished at once with the values a
        // walk would give it (background, t = inf, prim = -1, no counts) and its lane takes the next
        // pixel in another pass, so the lanes walk rays that can hit something.
        for (int pass = 0; pass < kRefillPasses; ++pass) {
            const uint64_t idle = __ballot(!busy);
            const uint32_t n_idle = (uint32_t)__popcll(idle);
            if (exhausted || !(n_idle >= kRefillIdle || n_idle == 64u)) break;
            if (tile_px == 64u) {
                if (++tile >= pool_end) {
                    uint32_t u0 = 0;
                    if (lane == 0) u0 = atomicAdd(&ws[0], grab);
                    u0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)u0);
                    if (u0 >= total) {
                        exhausted = true;
                    } else {
                        tile = u0;
                        pool_end = u0 + grab < total ? u0 + grab : total;
                    }
                }
                if (!exhausted) {
                    tile_px = 0;
                    uint32_t lo = 0, hi = n_jobs;  // largest j with pref[j] <= tile
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (pref[mid] <= tile) lo = mid; else hi = mid;
                    }
                    tile_job = (uint32_t)__builtin_amdgcn_readfirstlane((int)lo);
                }
            }
            if (!exhausted && tile_job != job && n_idle == 64u) {  // drained: move to the tile's job
                flush();
                job = tile_job;
                const rt_render_job& jb = jobs[job];
                sc = job_scene(jb);
                cam = jb.cam;
                p = job_params(jb);
                out = job_out(jb);
                tiles_x = (p.width + 7u) / 8u;
                job_first = pref[job];
                // the scene box (written with the hierarchy's root node, so only when the root is
                // an internal node) contains every child box of the root
                cull = sc.root >= 0 && sc.root != RT_EMPTY_ROOT && jb.scene.bounds != nullptr;
                if (cull) {
                    const float* bb = as_global((const float*)jb.scene.bounds);
                    blo0 = bb[0]; blo1 = bb[1]; blo2 = bb[2];
                    bhi0 = bb[4]; bhi1 = bb[5]; bhi2 = bb[6];
                }
            }
            bool culled = false;
            if (!exhausted && tile_job == job) {
                const uint32_t take = 64u - tile_px < n_idle ? 64u - tile_px : n_idle;
                if (!busy) {
                    const uint32_t k = lane_rank(idle);
                    if (k < take) {
                        const uint32_t q = tile_px + k, local = tile - job_first;
                        TilePixel px;
                        px.col = (local % tiles_x) * 8u + (q & 7u);
                        px.row = (local / tiles_x) * 8u + (q >> 3);
                        px.valid = px.col < p.width && px.row < p.height;
                        if (px.valid) {
                            col = px.col;
                            row = px.row;
                            r = camera_ray(cam, p.width, p.height, px);
                            h.t = r.tmax;
                            h.u = h.v = 0.0f;
                            h.prim = h.slot = -1;
                            walk_start(r, sc.root, w);
                            skip = -1;
                            phase = 0;
                            if (cull) {  // the node test of the root's children, on their union
                                const float x0 = fmaf(blo0, w.idx, -w.oxx), x1 = fmaf(bhi0, w.idx, -w.oxx);
                                const float y0 = fmaf(blo1, w.idy, -w.oyy), y1 = fmaf(bhi1, w.idy, -w.oyy);
                                const float z0 = fmaf(blo2, w.idz, -w.ozz), z1 = fmaf(bhi2, w.idz, -w.ozz);
                                const float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), r.tmin));
                                const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1)) * kTfarScale;
                                culled = !(tn <= fminf(tf, h.t));
                            }
                            if (culled) {
                                const uint64_t pix = (uint64_t)row * p.width + col;
                                if (out.t) out.t[pix] = INFINITY;
                                if (out.prim) out.prim[pix] = -1;
                                out.rgb[3 * pix] = p.background[0];
                                out.rgb[3 * pix + 1] = p.background[1];
                                out.rgb[3 * pix + 2] = p.background[2];
                            } else {
                                busy = true;
                            }
                        }
                    }
                }
                tile_px += take;
            }
            if (__ballot(culled) == 0) break;  // no lane freed: as full as this refill makes them
        }
        if (__ballot(busy) == 0) {
            if (exhausted) break;
            continue;
        }

        // ---- walk
        int st = kWalkOn;
        if (busy) {
            if (sc.root == RT_EMPTY_ROOT)
                st = kWalkDone;
            else
                st = lane4_steps<false, RT_PARK_RENDER>(sc.nodes4, sc.tris, r, skip, h, w, stk, 64, lane_stack, kRefillSteps,
                                        phase == 1u);
        }
        if (__ballot(st == kWalkOverflow)) {  // rare: the whole wave traces those rays as a packet
            const bool redo0 = st == kWalkOverflow && phase == 0u, redo1 = st == kWalkOverflow && phase == 1u;
            if (__ballot(redo0)) {
                Hit h2;
                traverse_packet<false>(sc.nodes4, sc.tris, sc.root, r, redo0, -1, h2, pk);
                if (redo0) h = h2;
            }
            if (__ballot(redo1)) {
                Hit h2;
                traverse_packet<true>(sc.nodes4, sc.tris, sc.root, r, redo1, skip, h2, pk);
                if (redo1) h = h2;
            }
            if (st == kWalkOverflow) st = kWalkDone;
        }

        // ---- finish: shade after the primary walk (and start the shadow walk), store after the last
        const bool fin0 = st == kWalkDone && phase == 0u, fin1 = st == kWalkDone && phase == 1u;
        bool hit = false, go_shadow = false, occluded = false, store = false;
        if (fin0) {
            hit = h.prim >= 0;
            ndl = 0.0f;
            alb0 = alb1 = alb2 = 1.0f;
            if (hit) {
                refine_hit(sc.tris, r, h);
                float tu = 0.0f, tv = 0.0f;
                int32_t mesh = 0;
                shade_terms(sc, p, h, ndl, tu, tv, mesh);
                float alb[3] = {1.0f, 1.0f, 1.0f};
                albedo(p, mesh, tu, tv, alb);
                alb0 = alb[0];
                alb1 = alb[1];
                alb2 = alb[2];
            }
            const uint64_t pix = (uint64_t)row * p.width + col;
            if (out.t) out.t[pix] = hit ? h.t : INFINITY;
            if (out.prim) out.prim[pix] = hit ? h.prim : -1;
            go_shadow = hit && (p.flags & RT_RENDER_SHADOWS) && ndl > 0.0f;
            if (go_shadow) {
                r = shadow_ray(p, r, h);
                skip = h.slot;
                h.t = r.tmax;
                h.u = h.v = 0.0f;
                h.prim = h.slot = -1;
                walk_start(r, sc.root, w);
                phase = 1;
            } else {
                store = true;
            }
        }
        if (fin1) {
            occluded = h.prim >= 0;
            hit = true;
            store = true;
        }
        if (store) {
            const uint64_t pix = (uint64_t)row * p.width + col;
            if (hit) {
                const float kk = fmaf(occluded ? 0.0f : 1.0f, ndl, p.ambient);
                out.rgb[3 * pix] = alb0 * kk;
                out.rgb[3 * pix + 1] = alb1 * kk;
                out.rgb[3 * pix + 2] = alb2 * kk;
            } else {
                out.rgb[3 * pix] = p.background[0];
                out.rgb[3 * pix + 1] = p.background[1];
                out.rgb[3 * pix + 2] = p.background[2];
            }
            busy = false;
        }
        n_hit += (uint32_t)__popcll(__ballot(fin0 && hit));
        n_sh += (uint32_t)__popcll(__ballot(go_shadow));
        n_oc += (uint32_t)__popcll(__ballot(occluded));
    }
    flush();
}

__global__ void k_zero_u32(uint32_t* c, uint32_t n) {
    if (threadIdx.x < n) c[threadIdx.x] = 0;
}

__global__ void k_zero_words(uint32_t* w, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) w[i] = 0;
}

// stats[0..2] = the sums of the kCountSlots partial counts (stats zeroed before)
__global__ void __launch_bounds__(256) k_sum_counts(const uint32_t* __restrict__ part, uint32_t* __restrict__ stats) {
    __shared__ uint32_t sum[3];
    if (threadIdx.x < 3) sum[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kCountSlots; i += blockDim.x)
        for (int c = 0; c < 3; ++c)
            if (part[i * kCountStride + c]) atomicAdd(&sum[c], part[i * kCountStride + c]);
    __syncthreads();
    if (threadIdx.x < 3) stats[threadIdx.x] = sum[threadIdx.x];
}

int check_rays(const rt_rays* r, const char* what) {
    if (!r || !r->ox || !r->oy || !r->oz || !r->dx || !r->dy || !r->dz)
        return set_error(RT_EINVAL, std::string(what) + ": ray field pointer is NULL");
    if (r->stride < 4 || (r->stride & 3))
        return set_error(RT_EINVAL, std::string(what) + ": ray stride must be a multiple of 4");
    return RT_OK;
}

inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }
// Blocks of a k_query launch: as many as the device holds at once (the CU count times the
// kernel's residency per CU at its LDS size, cached per device, kernel variant and stack size), or
// fewer when the buffer has fewer 64-ray groups.
uint32_t query_blocks(const void* kern, int variant, uint32_t lane_stack, uint64_t n) {
    static std::mutex mu;
    static int per_cu_cache[64][4][RT_LANE_STACK_MAX + 1] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int per_cu = 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        int& slot = per_cu_cache[dev][variant][lane_stack];
        if (slot == 0) {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 256, 4u * 256u * lane_stack) !=
                    hipSuccess || nb <= 0)
                nb = 1;
            slot = nb;
        }
        per_cu = slot;
    }
    const uint64_t groups = (n + 63) / 64, waves = (uint64_t)per_cu * device_cu_count() * 4u;
    const uint64_t w = groups < waves ? groups : waves;
    return (uint32_t)((w + 3) / 4);
}

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" int rt_trace_closest(const rt_scene* scene, const rt_rays* rays, uint32_t n,
                                const rt_hits* hits, const rt_tuning* tuning, void* stream) {
    int e = check_scene(scene, "rt_trace_closest");
    if (e) return e;
    if ((e = check_rays(rays, "rt_trace_closest"))) return e;
    if (!hits || hits->stride < 4 || (hits->stride & 3))
        return set_error(RT_EINVAL, "rt_trace_closest: bad hits descriptor");
    rt_tuning tu;
    if ((e = resolve_tuning(tuning, tu, "rt_trace_closest"))) return e;
    if (n == 0) return RT_OK;
    const size_t lds = 4u * 256u * tu.lane_stack;
    auto k = tu.query == RT_TRAVERSAL_LANE ? k_query<false, RT_TRAVERSAL_LANE>
                                           : k_query<false, RT_TRAVERSAL_BINARY>;
    const uint32_t blocks = query_blocks((const void*)k, tu.query == RT_TRAVERSAL_LANE, tu.lane_stack, n);
    hipLaunchKernelGGL(k, dim3(blocks), dim3(256), lds, (hipStream_t)stream, scene_args(scene), *rays, n,
                       *hits, (uint8_t*)nullptr, (int)tu.lane_stack, 4u * blocks);
    return check_launch("rt_trace_closest");
}

extern "C" int rt_trace_any(const rt_scene* scene, const rt_rays* rays, uint32_t n,
                            uint8_t* occluded, const rt_tuning* tuning, void* stream) {
    int e = check_scene(scene, "rt_trace_any");
    if (e) return e;
    if ((e = check_rays(rays, "rt_trace_any"))) return e;
    if (!occluded) return set_error(RT_EINVAL, "rt_trace_any: occluded is NULL");
    rt_tuning tu;
    if ((e = resolve_tuning(tuning, tu, "rt_trace_any"))) return e;
    if (n == 0) return RT_OK;
    const size_t lds = 4u * 256u * tu.lane_stack;
    auto k = tu.query == RT_TRAVERSAL_LANE ? k_query<true, RT_TRAVERSAL_LANE>
                                           : k_query<true, RT_TRAVERSAL_BINARY>;
    const uint32_t blocks =
        query_blocks((const void*)k, 2 + (tu.query == RT_TRAVERSAL_LANE), tu.lane_stack, n);
    hipLaunchKernelGGL(k, dim3(blocks), dim3(256), lds, (hipStream_t)stream, scene_args(scene), *rays, n,
                       rt_hits{}, occluded, (int)tu.lane_stack, 4u * blocks);
    return check_launch("rt_trace_any");
}

extern "C" int rt_generate_rays(const rt_camera* cam, uint32_t W, uint32_t H, float* ox, float* oy,
                                float* oz, float* dx, float* dy, float* dz, void* stream) {
    if (!cam || !ox || !oy || !oz || !dx || !dy || !dz)
        return set_error(RT_EINVAL, "rt_generate_rays: NULL argument");
    const uint64_t n = (uint64_t)W * H;
    if (n == 0) return RT_OK;
    if (n > 0xFFFFFFFFull) return set_error(RT_ECAPACITY, "rt_generate_rays: image too large");
    hipLaunchKernelGGL(k_generate_rays, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       *cam, W, H, ox, oy, oz, dx, dy, dz);
    return check_launch("rt_generate_rays");
}

extern "C" uint64_t rt_trace_camera_workspace_bytes(uint32_t width, uint32_t height) {
    (void)width;
    (void)height;
    return 0;
}

extern "C" int rt_trace_camera(const rt_scene* scene, const rt_camera* cam, uint32_t W, uint32_t H,
                               const rt_hits* hits, void* workspace, uint64_t ws_bytes,
                               const rt_tuning* tuning, void* stream) {
    (void)workspace;
    (void)ws_bytes;
    int e = check_scene(scene, "rt_trace_camera");
    if (e) return e;
    if (!cam || !hits || hits->stride < 4 || (hits->stride & 3))
        return set_error(RT_EINVAL, "rt_trace_camera: bad camera or hits descriptor");
    const uint64_t n = (uint64_t)W * H;
    if (n == 0) return RT_OK;
    if (n > 0x7FFFFFFFull) return set_error(RT_ECAPACITY, "rt_trace_camera: image too large");
    rt_tuning tu;
    if ((e = resolve_tuning(tuning, tu, "rt_trace_camera"))) return e;
    hipStream_t s = (hipStream_t)stream;
    if (tu.primary == RT_TRAVERSAL_QUAD) {  // 8x8-pixel blocks
        const TileMap tm8 = make_tile_map(W, H, tu.super_tile, 8);
        hipLaunchKernelGGL(k_trace_camera_quad, dim3(32u * ((tile_map_blocks(tm8) + 7u) / 8u)), dim3(64), 0, s,
                           scene_args(scene), *cam, W, H, tm8, *hits);
        return check_launch("rt_trace_camera");
    }
    const TileMap tm = make_tile_map(W, H, tu.super_tile);
    using K = decltype(&k_trace_camera<0>);
    static const K kerns[3] = {k_trace_camera<0>, k_trace_camera<1>, k_trace_camera<2>};
    const int pm = resolve_primary(tu.primary, *cam, W, H);
    hipLaunchKernelGGL(kerns[pm], dim3(tile_map_blocks(tm)), dim3(256), 0, s, scene_args(scene), *cam,
                       W, H, tm, (int)tu.lane_stack, (int)tu.packet_budget, *hits);
    return check_launch("rt_trace_camera");
}

extern "C" uint64_t rt_render_workspace_bytes(uint32_t n_pixels) {
    (void)n_pixels;
    return (uint64_t)kCountSlots * kCountStride * sizeof(uint32_t);
}

extern "C" int rt_render(const rt_scene* scene, const rt_camera* cam, const rt_render_params* p,
                         const rt_render_out* out, void* workspace, uint64_t ws_bytes,
                         const rt_tuning* tuning, void* stream) {
    int e = check_scene(scene, "rt_render");
    if (e) return e;
    if (!cam || !p || !out) return set_error(RT_EINVAL, "rt_render: NULL argument");
    const uint64_t n = (uint64_t)p->width * p->height;
    if (n == 0) return RT_OK;
    if (!out->rgb) return set_error(RT_EINVAL, "rt_render: rgb output is NULL");
    if (n > 0x7FFFFFFFull) return set_error(RT_ECAPACITY, "rt_render: image too large");
    if (p->n_textures && !p->textures_dev)
        return set_error(RT_EINVAL, "rt_render: n_textures > 0 but textures_dev is NULL");
    rt_tuning tu;
    if ((e = resolve_tuning(tuning, tu, "rt_render"))) return e;
    hipStream_t s = (hipStream_t)stream;
    if (out->stats) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, out->stats, 4u);
    const TileMap tm = make_tile_map(p->width, p->height, tu.super_tile);
    using K = decltype(&k_render<0, 0>);
    static const K kerns[3][3] = {{k_render<0, 0>, k_render<0, 1>, k_render<0, 2>},
                                  {k_render<1, 0>, k_render<1, 1>, k_render<1, 2>},
                                  {k_render<2, 0>, k_render<2, 1>, k_render<2, 2>}};
    // QUAD: the quad-lane traversal for both rays (16-ray packets, 4 lanes per ray), at every frame
    // size: at 1920x1080 on the 1.31M-triangle scene 1.19 ms against 1.67 ms for packet primary +
    // lane shadows and 1.45 ms for lane + lane, and at least as fast on smaller scenes
    if (tu.primary == RT_TRAVERSAL_QUAD) {
        const uint32_t tx8 = (p->width + 7) / 8, ty8 = (p->height + 7) / 8;
        const bool partial = out->stats && workspace && ws_bytes >= rt_render_workspace_bytes((uint32_t)n);
        if (out->stats && !partial) {
            hipLaunchKernelGGL(k_render_quad<false>, dim3(tx8 * ty8), dim3(256), 0, s, scene_args(scene), *cam, *p,
                               tx8, out->rgb, out->t, out->prim, out->stats);
        } else {
            uint32_t* part = partial ? (uint32_t*)workspace : nullptr;
            if (part) hipLaunchKernelGGL(k_zero_words, dim3(8), dim3(256), 0, s, part, kCountSlots * kCountStride);
            hipLaunchKernelGGL(k_render_quad<true>, dim3(32u * ((tx8 * ty8 + 7u) / 8u)), dim3(64), 0, s,
                               scene_args(scene), *cam, *p, tx8, out->rgb, out->t, out->prim, part);
            if (part) hipLaunchKernelGGL(k_sum_counts, dim3(1), dim3(256), 0, s, part, out->stats);
        }
        return check_launch("rt_render");
    }
    const int pm = resolve_primary(tu.primary, *cam, p->width, p->height);
    hipLaunchKernelGGL(kerns[pm][tu.shadow], dim3(tile_map_blocks(tm)), dim3(256), 0, s,
                       scene_args(scene), *cam, *p, tm, (int)tu.lane_stack, (int)tu.packet_budget,
                       out->rgb, out->t, out->prim, out->stats);
    return check_launch("rt_render");
}

extern "C" uint64_t rt_render_batch_workspace_bytes(uint32_t n_jobs) {
    return 4ull * (kBatchPrefixWord + (uint64_t)n_jobs + 1);
}

extern "C" const uint32_t* rt_render_batch_status(const void* workspace_dev) {
    return workspace_dev ? (const uint32_t*)workspace_dev + 1 : nullptr;
}

namespace rt {
namespace {

// Validation shared by the batch entry points; RT_OK or an error code with its message.
int check_batch_args(const rt_render_job* jobs_dev, uint32_t n_jobs, const rt_batch_config* cfg,
                     const char* what) {
    if (!jobs_dev || !cfg) return set_error(RT_EINVAL, std::string(what) + ": NULL argument");
    if (n_jobs > RT_BATCH_MAX_JOBS)
        return set_error(RT_ECAPACITY, std::string(what) + ": more than RT_BATCH_MAX_JOBS jobs");
    auto alg_ok = [](int32_t a) { return a >= RT_TRAVERSAL_PACKET && a <= RT_TRAVERSAL_HYBRID; };
    if (cfg->waves_per_cu < 1 || cfg->waves_per_cu > 32 || cfg->tiles_per_grab < 1 ||
        cfg->tiles_per_grab > 64 || cfg->lane_stack < 1 || cfg->lane_stack > RT_LANE_STACK_MAX ||
        !(alg_ok(cfg->primary) || cfg->primary == RT_TRAVERSAL_AUTO || cfg->primary == RT_TRAVERSAL_QUAD) ||
        !alg_ok(cfg->shadow) ||
        cfg->packet_budget < 1)
        return set_error(RT_EINVAL, std::string(what) + ": rt_batch_config field out of range");
    return RT_OK;
}

// The two launches of a batch (prepare + persistent kernel) on `ws`, a workspace of
// rt_render_batch_workspace_bytes(n_jobs) bytes whose every word the prepare kernel rewrites.
int launch_batch(const rt_render_job* jobs_dev, uint32_t n_jobs, const rt_batch_config* cfg,
                 uint32_t* ws, hipStream_t s, const char* what) {
    const bool quad = cfg->primary == RT_TRAVERSAL_QUAD;
    hipLaunchKernelGGL(k_batch_prepare, dim3(1), dim3(1024), 0, s, jobs_dev, n_jobs, quad ? 4u : 8u, ws);
    const bool lds_stack = !quad && (cfg->primary != RT_TRAVERSAL_PACKET || cfg->shadow != RT_TRAVERSAL_PACKET);
    const size_t lds = 4u * ((n_jobs + 1 + 3) & ~3u) + (lds_stack ? 4u * 64u * kLaneStack : 0u);
    const uint32_t grid = (uint32_t)device_cu_count() * cfg->waves_per_cu;
    using K = decltype(&k_render_batch<0, 0>);
    static const K kerns[4][3] = {
        {k_render_batch<0, 0>, k_render_batch<0, 1>, k_render_batch<0, 2>},
        {k_render_batch<1, 0>, k_render_batch<1, 1>, k_render_batch<1, 2>},
        {k_render_batch<2, 0>, k_render_batch<2, 1>, k_render_batch<2, 2>},
        {k_render_batch<3, 0>, k_render_batch<3, 1>, k_render_batch<3, 2>}};
    if (cfg->primary == RT_TRAVERSAL_LANE && cfg->shadow == RT_TRAVERSAL_LANE) {
        hipLaunchKernelGGL(k_render_batch_refill, dim3(grid), dim3(64), lds, s, jobs_dev, n_jobs, ws,
                           cfg->tiles_per_grab, (int)cfg->lane_stack);
        return check_launch(what);
    }
    const K kern = quad ? k_render_batch<RT_TRAVERSAL_QUAD, RT_TRAVERSAL_QUAD> : kerns[cfg->primary][cfg->shadow];
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64), lds, s, jobs_dev, n_jobs, ws, cfg->tiles_per_grab,
                       (int)cfg->lane_stack, (int)cfg->packet_budget);
    return check_launch(what);
}

}  // namespace
}  // namespace rt

extern "C" int rt_render_batch(const rt_render_job* jobs_dev, uint32_t n_jobs, const rt_batch_config* cfg,
                               void* workspace, uint64_t ws_bytes, void* stream) {
    int e = check_batch_args(jobs_dev, n_jobs, cfg, "rt_render_batch");
    if (e) return e;
    if (n_jobs == 0) return RT_OK;
    if (!workspace || ws_bytes < rt_render_batch_workspace_bytes(n_jobs))
        return set_error(RT_EINVAL, "rt_render_batch: workspace too small");
    if ((uintptr_t)workspace & 15) return set_error(RT_EINVAL, "rt_render_batch: workspace must be 16-byte aligned");
    return launch_batch(jobs_dev, n_jobs, cfg, (uint32_t*)workspace, (hipStream_t)stream, "rt_render_batch");
}

// ---- allocation-free batched render ----
// The batch workspace of RT_BATCH_MAX_JOBS jobs as a file-scope device array: every device the
// library runs on holds its own copy until the process ends. In host code `scratch` only names the
// symbol; its device address comes from hipGetSymbolAddress, looked up once per device (with that
// device current) and kept in g_scratch.
constexpr uint32_t kStaticBatchWords = kBatchPrefixWord + RT_BATCH_MAX_JOBS + 1;  // 16,644 bytes
static __device__ __align__(256) uint32_t scratch[kStaticBatchWords];

namespace rt {
namespace {
std::mutex g_scratch_mutex;
uint32_t* g_scratch[64] = {nullptr};

int scratch_address(uint32_t** out, const char* what) {
    int dev = 0;
    if (int e = check_hip(hipGetDevice(&dev), what)) return e;
    if (dev < 0 || dev >= 64) return set_error(RT_ECAPACITY, std::string(what) + ": device id above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        if (int e = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), what)) return e;
        g_scratch[dev] = (uint32_t*)p;
    }
    *out = g_scratch[dev];
    return RT_OK;
}
}  // namespace
}  // namespace rt

extern "C" int rt_render_batch_static(const rt_render_job* jobs_dev, uint32_t n_jobs,
                                      const rt_batch_config* cfg, void* stream) {
    int e = check_batch_args(jobs_dev, n_jobs, cfg, "rt_render_batch_static");
    if (e) return e;
    if (n_jobs == 0) return RT_OK;
    uint32_t* ws = nullptr;
    if ((e = scratch_address(&ws, "rt_render_batch_static"))) return e;
    return launch_batch(jobs_dev, n_jobs, cfg, ws, (hipStream_t)stream, "rt_render_batch_static");
}
