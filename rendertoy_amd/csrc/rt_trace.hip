// rt_trace.hip — closest-hit / any-hit queries and the two-kernel render path.
//
// Queries (rendering/_raycaster.py:35-36, Raycaster.ray_cast — a stub in the reference):
//   k_trace_closest / k_trace_any: one lane per ray, rays and hits read/written through
//   (pointer, byte stride) fields so SoA (stride 4) and the AoS Ray/Hit records both coalesce.
// Render:
//   k_render_primary: each wave64 owns an 8x8 pixel tile (a block = 16x16 pixels) so the lanes of
//     a wave start with nearly identical traversals; rays are generated in registers, traced, and
//     the hit lanes are appended to a compacted hit queue (ballot + mbcnt, one atomic per wave).
//     Miss pixels get the background colour directly.
//   k_shade: persistent grid-stride pass over the compacted queue — fused Lambert, shadow ray
//     (any-hit, skipping the hit's own triangle) and RGBA8 texture fetch; writes RGB float32.
#include <algorithm>

#include "rt_common.h"

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

struct SceneArgs {
    const float4* nodes;
    const float4* tris;
    const float4* shade;
    const int32_t* mesh_offsets;
    int32_t root;
    uint32_t n_meshes;
};

__global__ void __launch_bounds__(256) k_trace_closest(SceneArgs sc, rt_rays rays, uint32_t n,
                                                        rt_hits hits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Ray r;
    r.ox = ld_f(rays.ox, i, rays.stride);
    r.oy = ld_f(rays.oy, i, rays.stride);
    r.oz = ld_f(rays.oz, i, rays.stride);
    r.dx = ld_f(rays.dx, i, rays.stride);
    r.dy = ld_f(rays.dy, i, rays.stride);
    r.dz = ld_f(rays.dz, i, rays.stride);
    r.tmin = rays.tmin ? ld_f(rays.tmin, i, rays.stride) : rays.tmin_default;
    r.tmax = rays.tmax ? ld_f(rays.tmax, i, rays.stride) : rays.tmax_default;
    Hit h;
    traverse<false>(sc.nodes, sc.tris, sc.root, r, -1, h);
    if (h.prim >= 0) refine_hit(sc.tris, r, h);
    const bool hit = h.prim >= 0;
    int32_t mesh = -1, index = -1;
    if (hit) {
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

__global__ void __launch_bounds__(256) k_trace_any(SceneArgs sc, rt_rays rays, uint32_t n,
                                                    uint8_t* __restrict__ occluded) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Ray r;
    r.ox = ld_f(rays.ox, i, rays.stride);
    r.oy = ld_f(rays.oy, i, rays.stride);
    r.oz = ld_f(rays.oz, i, rays.stride);
    r.dx = ld_f(rays.dx, i, rays.stride);
    r.dy = ld_f(rays.dy, i, rays.stride);
    r.dz = ld_f(rays.dz, i, rays.stride);
    r.tmin = rays.tmin ? ld_f(rays.tmin, i, rays.stride) : rays.tmin_default;
    r.tmax = rays.tmax ? ld_f(rays.tmax, i, rays.stride) : rays.tmax_default;
    Hit h;
    traverse<true>(sc.nodes, sc.tris, sc.root, r, -1, h);
    occluded[i] = h.prim >= 0 ? 1 : 0;
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

// Render workspace: counters (256 B) then the SoA hit queue.
struct Queue {
    uint32_t* counter;  // [0] = queue length
    uint32_t* pix;
    uint32_t* slot;
    float* u;
    float* v;
    float* t;
};

inline uint64_t a256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

Queue carve_queue(char* base, uint64_t n) {
    Queue q{};
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        char* p = base ? base + off : nullptr;
        off = a256(off + bytes);
        return p;
    };
    q.counter = (uint32_t*)take(256);
    q.pix = (uint32_t*)take(4 * n);
    q.slot = (uint32_t*)take(4 * n);
    q.u = (float*)take(4 * n);
    q.v = (float*)take(4 * n);
    q.t = (float*)take(4 * n);
    return q;
}

uint64_t queue_bytes(uint64_t n) {
    uint64_t off = a256(256);
    off += 5 * a256(4 * n);
    return off;
}

__global__ void k_zero_u32(uint32_t* c, uint32_t n) {
    if (threadIdx.x < n) c[threadIdx.x] = 0;
}

constexpr int kTile = 16;  // block = 16x16 pixels, wave = 8x8

__global__ void __launch_bounds__(256) k_render_primary(SceneArgs sc, rt_camera cam, uint32_t W,
                                                         uint32_t H, uint32_t tiles_x,
                                                         float bg0, float bg1, float bg2,
                                                         float* __restrict__ rgb,
                                                         float* __restrict__ out_t,
                                                         int32_t* __restrict__ out_prim, Queue q) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const uint32_t col = tx * kTile + (wave & 1) * 8 + (lane & 7);
    const uint32_t row = ty * kTile + (wave >> 1) * 8 + (lane >> 3);
    const bool valid = col < W && row < H;
    Hit h;
    h.prim = -1;
    if (valid) {
        Ray r;
        r.ox = cam.eye[0];
        r.oy = cam.eye[1];
        r.oz = cam.eye[2];
        camera_dir(cam, W, H, col, row, r.dx, r.dy, r.dz);
        r.tmin = 0.0f;
        r.tmax = INFINITY;
        traverse<false>(sc.nodes, sc.tris, sc.root, r, -1, h);
        if (h.prim >= 0) refine_hit(sc.tris, r, h);
    }
    const bool hit = valid && h.prim >= 0;
    const uint64_t pix = (uint64_t)row * W + col;
    if (valid) {
        if (out_t) out_t[pix] = hit ? h.t : INFINITY;
        if (out_prim) out_prim[pix] = h.prim;
        if (!hit) {
            rgb[3 * pix] = bg0;
            rgb[3 * pix + 1] = bg1;
            rgb[3 * pix + 2] = bg2;
        }
    }
    const uint32_t k = wave_append(q.counter, hit);
    if (hit) {
        q.pix[k] = (uint32_t)pix;
        q.slot[k] = (uint32_t)h.slot;
        q.u[k] = h.u;
        q.v[k] = h.v;
        q.t[k] = h.t;
    }
}

// Camera closest hit: ray generation fused into the traversal, one hit record per pixel.
__global__ void __launch_bounds__(256) k_trace_camera(SceneArgs sc, rt_camera cam, uint32_t W,
                                                       uint32_t H, uint32_t tiles_x, rt_hits hits) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const uint32_t col = tx * kTile + (wave & 1) * 8 + (lane & 7);
    const uint32_t row = ty * kTile + (wave >> 1) * 8 + (lane >> 3);
    if (col >= W || row >= H) return;
    Ray r;
    r.ox = cam.eye[0];
    r.oy = cam.eye[1];
    r.oz = cam.eye[2];
    camera_dir(cam, W, H, col, row, r.dx, r.dy, r.dz);
    r.tmin = 0.0f;
    r.tmax = INFINITY;
    Hit h;
    traverse<false>(sc.nodes, sc.tris, sc.root, r, -1, h);
    if (h.prim >= 0) refine_hit(sc.tris, r, h);
    const bool hit = h.prim >= 0;
    const uint64_t i = (uint64_t)row * W + col;
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

RT_DEV int wrap_i(int i, int n) {
    const int r = i % n;
    return r < 0 ? r + n : r;
}

RT_DEV float4 texel(const uint8_t* __restrict__ tex, int W, int x, int y) {
    const uint32_t p = *(const uint32_t*)(tex + ((uint64_t)y * W + x) * 4);
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

RT_DEV uint32_t wave_sum(uint32_t x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    return x;
}

__global__ void __launch_bounds__(256) k_shade(SceneArgs sc, rt_camera cam, uint32_t W,
                                                uint32_t H, rt_render_params p,
                                                float* __restrict__ rgb, Queue q,
                                                uint32_t* __restrict__ stats) {
    const uint32_t count = q.counter[0];
    uint32_t n_shadow = 0, n_occl = 0;
    const bool shadows = (p.flags & RT_RENDER_SHADOWS) != 0;
    const bool bilinear = (p.flags & RT_RENDER_BILINEAR) != 0;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
        const uint32_t pix = q.pix[k];
        const int32_t slot = (int32_t)q.slot[k];
        const float u = q.u[k], v = q.v[k], t = q.t[k];
        const float4* s = sc.shade + (uint64_t)slot * 4;
        const float4 s0 = s[0], s1 = s[1], s2 = s[2], s3 = s[3];
        const float w0 = 1.0f - u - v;
        const float nx = fmaf(w0, s0.x, fmaf(u, s0.w, v * s1.z));
        const float ny = fmaf(w0, s0.y, fmaf(u, s1.x, v * s1.w));
        const float nz = fmaf(w0, s0.z, fmaf(u, s1.y, v * s2.x));
        const float l2 = fmaf(nx, nx, fmaf(ny, ny, nz * nz));
        const float il = l2 > 0.0f ? 1.0f / sqrtf(l2) : 0.0f;
        float ndl = fmaf(nx * il, p.light[0], fmaf(ny * il, p.light[1], (nz * il) * p.light[2]));
        ndl = fmaxf(ndl, 0.0f);
        float vis = 1.0f;
        if (shadows && ndl > 0.0f) {
            float dx, dy, dz;
            camera_dir(cam, W, H, pix % W, pix / W, dx, dy, dz);
            Ray r;
            r.ox = fmaf(t, dx, cam.eye[0]);
            r.oy = fmaf(t, dy, cam.eye[1]);
            r.oz = fmaf(t, dz, cam.eye[2]);
            r.dx = p.light[0];
            r.dy = p.light[1];
            r.dz = p.light[2];
            r.tmin = p.shadow_tmin;
            r.tmax = INFINITY;
            Hit sh;
            traverse<true>(sc.nodes, sc.tris, sc.root, r, slot, sh);
            ++n_shadow;
            if (sh.prim >= 0) {
                vis = 0.0f;
                ++n_occl;
            }
        }
        float alb[3] = {1.0f, 1.0f, 1.0f};
        const int32_t mesh = __float_as_int(s3.w);
        const int32_t ti = p.mesh_texture_dev ? p.mesh_texture_dev[mesh] : -1;
        if (ti >= 0 && (uint32_t)ti < p.n_textures) {
            const float tu = fmaf(w0, s2.y, fmaf(u, s2.w, v * s3.y));
            const float tv = fmaf(w0, s2.z, fmaf(u, s3.x, v * s3.z));
            tex_sample(p.textures_dev[ti], tu, tv, bilinear, alb);
        }
        const float kk = fmaf(vis, ndl, p.ambient);
        rgb[3 * (uint64_t)pix] = alb[0] * kk;
        rgb[3 * (uint64_t)pix + 1] = alb[1] * kk;
        rgb[3 * (uint64_t)pix + 2] = alb[2] * kk;
    }
    if (stats) {
        const uint32_t a = wave_sum(n_shadow), b = wave_sum(n_occl);
        if ((threadIdx.x & 63) == 0) {
            if (a) atomicAdd(&stats[1], a);
            if (b) atomicAdd(&stats[2], b);
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) stats[0] = count;
    }
}

SceneArgs scene_args(const rt_scene* s) {
    SceneArgs a;
    a.nodes = (const float4*)s->nodes;
    a.tris = (const float4*)s->tris;
    a.shade = (const float4*)s->shade;
    a.mesh_offsets = s->mesh_offsets;
    a.root = s->root;
    a.n_meshes = s->n_meshes;
    return a;
}

int check_scene(const rt_scene* s, const char* what) {
    if (!s) return set_error(RT_EINVAL, std::string(what) + ": scene is NULL");
    if (s->root != RT_EMPTY_ROOT && (!s->nodes || !s->tris || !s->mesh_offsets))
        return set_error(RT_EINVAL, std::string(what) + ": scene arrays are NULL");
    if (s->root != RT_EMPTY_ROOT && s->n_tris == 0)
        return set_error(RT_EINVAL, std::string(what) + ": scene root set but n_tris is 0");
    return RT_OK;
}

int check_rays(const rt_rays* r, const char* what) {
    if (!r || !r->ox || !r->oy || !r->oz || !r->dx || !r->dy || !r->dz)
        return set_error(RT_EINVAL, std::string(what) + ": ray field pointer is NULL");
    if (r->stride < 4 || (r->stride & 3))
        return set_error(RT_EINVAL, std::string(what) + ": ray stride must be a multiple of 4");
    return RT_OK;
}

inline uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" int rt_trace_closest(const rt_scene* scene, const rt_rays* rays, uint32_t n,
                                const rt_hits* hits, void* stream) {
    int e = check_scene(scene, "rt_trace_closest");
    if (e) return e;
    if ((e = check_rays(rays, "rt_trace_closest"))) return e;
    if (!hits || hits->stride < 4 || (hits->stride & 3))
        return set_error(RT_EINVAL, "rt_trace_closest: bad hits descriptor");
    if (n == 0) return RT_OK;
    hipLaunchKernelGGL(k_trace_closest, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       scene_args(scene), *rays, n, *hits);
    return check_launch("rt_trace_closest");
}

extern "C" int rt_trace_any(const rt_scene* scene, const rt_rays* rays, uint32_t n,
                            uint8_t* occluded, void* stream) {
    int e = check_scene(scene, "rt_trace_any");
    if (e) return e;
    if ((e = check_rays(rays, "rt_trace_any"))) return e;
    if (!occluded) return set_error(RT_EINVAL, "rt_trace_any: occluded is NULL");
    if (n == 0) return RT_OK;
    hipLaunchKernelGGL(k_trace_any, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       scene_args(scene), *rays, n, occluded);
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

extern "C" int rt_trace_camera(const rt_scene* scene, const rt_camera* cam, uint32_t W, uint32_t H,
                               const rt_hits* hits, void* stream) {
    int e = check_scene(scene, "rt_trace_camera");
    if (e) return e;
    if (!cam || !hits || hits->stride < 4 || (hits->stride & 3))
        return set_error(RT_EINVAL, "rt_trace_camera: bad camera or hits descriptor");
    const uint64_t n = (uint64_t)W * H;
    if (n == 0) return RT_OK;
    if (n > 0x7FFFFFFFull) return set_error(RT_ECAPACITY, "rt_trace_camera: image too large");
    const uint32_t tiles_x = (W + kTile - 1) / kTile, tiles_y = (H + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_trace_camera, dim3(tiles_x * tiles_y), dim3(256), 0, (hipStream_t)stream,
                       scene_args(scene), *cam, W, H, tiles_x, *hits);
    return check_launch("rt_trace_camera");
}

extern "C" uint64_t rt_render_workspace_bytes(uint32_t n_pixels) { return queue_bytes(n_pixels); }

extern "C" int rt_render(const rt_scene* scene, const rt_camera* cam, const rt_render_params* p,
                         const rt_render_out* out, void* workspace, uint64_t ws_bytes,
                         void* stream) {
    int e = check_scene(scene, "rt_render");
    if (e) return e;
    if (!cam || !p || !out || !out->rgb) return set_error(RT_EINVAL, "rt_render: NULL argument");
    const uint64_t n = (uint64_t)p->width * p->height;
    if (n == 0) return RT_OK;
    if (n > 0x7FFFFFFFull) return set_error(RT_ECAPACITY, "rt_render: image too large");
    if (p->n_textures && !p->textures_dev)
        return set_error(RT_EINVAL, "rt_render: n_textures > 0 but textures_dev is NULL");
    if (!workspace || ws_bytes < queue_bytes(n))
        return set_error(RT_EINVAL, "rt_render: workspace too small (need " +
                                        std::to_string(queue_bytes(n)) + " bytes)");
    if ((uintptr_t)workspace & 255) return set_error(RT_EINVAL, "rt_render: workspace must be 256-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    Queue q = carve_queue((char*)workspace, n);
    const SceneArgs sa = scene_args(scene);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, q.counter, 64u);
    if (out->stats) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, out->stats, 4u);
    const uint32_t tiles_x = (p->width + kTile - 1) / kTile, tiles_y = (p->height + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_render_primary, dim3(tiles_x * tiles_y), dim3(256), 0, s, sa, *cam, p->width,
                       p->height, tiles_x, p->background[0], p->background[1], p->background[2],
                       out->rgb, out->t, out->prim, q);
    const uint32_t grid = std::min<uint32_t>(blocks_for(n, 256), (uint32_t)device_cu_count() * 8);
    hipLaunchKernelGGL(k_shade, dim3(grid), dim3(256), 0, s, sa, *cam, p->width, p->height, *p,
                       out->rgb, q, out->stats);
    return check_launch("rt_render");
}
