// rt_paths.hip — incoherent rays: diffuse paths with seeded Philox sampling and ray compaction.
//
// No reference counterpart: the reference casts only primary rays (rendering/_raycaster.py:35-36,
// a stub). This is the divergent-ray workload of the engine: spp jittered camera rays per pixel,
// each continued by up to `bounces` cosine-weighted diffuse bounces, lit by a uniform sky.
//   sample radiance = sky * albedo^k if the path escapes after k bounces, 0 if it is still on a
//   surface after the last bounce; pixel = mean of its samples (fixed order).
// Kernels (all on the caller's stream, no host synchronisation, no memset: graph-capturable):
//   k_paths_init     counters
//   k_paths_primary  camera rays (QUAD: a 4x4 tile per one-wave workgroup; else an 8x8 tile per
//                    wave, AUTO coherence rule), hits append a bounce ray
//                    to queue 0 (wave ballot + one atomic per wave = compaction)
//   k_paths_bounce   per depth: the waves share the queue's 64-ray groups (count read on the
//                    device), per-lane walks with refill; each ray records the slot it hit
//   k_paths_emit     per depth, in queue order: misses write their sample, hits append their
//                    bounce to the other queue
//   k_paths_resolve  per-pixel mean, ray counts out
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), key = (seed, frame), counter =
// (pixel, sample, depth, 0); 32-bit outputs -> [0,1) as (x >> 8) * 2^-24.
#include <algorithm>

#include "rt_trace_shared.h"

namespace rt {
namespace {

RT_DEV uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1,
                       (uint32_t)p0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

RT_DEV float unit_float(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

// cos and sin of 2*pi*x for x in [0, 1): quadrant q = floor(4x), a = (4x - q) * pi/2 in [0, pi/2),
// Taylor polynomials of sin (to a^11) and cos (to a^12) by Horner with fmaf (truncation error below
// 6e-8 on the interval), then the quarter-turn rotation of the quadrant.
RT_DEV void cos_sin_2pi(float x, float& c, float& s) {
    const float x4 = x * 4.0f;
    const int q = (int)x4;
    const float a = (x4 - (float)q) * 1.57079632679489662f;
    const float a2 = a * a;
    float sp = fmaf(a2, -2.50521083854417188e-8f, 2.75573192239858907e-6f);
    sp = fmaf(a2, sp, -1.98412698412698413e-4f);
    sp = fmaf(a2, sp, 8.33333333333333333e-3f);
    sp = fmaf(a2, sp, -1.66666666666666667e-1f);
    sp = fmaf(a2, sp, 1.0f);
    const float sa = a * sp;
    float cp = fmaf(a2, 2.08767569878680990e-9f, -2.75573192239858907e-7f);
    cp = fmaf(a2, cp, 2.48015873015873016e-5f);
    cp = fmaf(a2, cp, -1.38888888888888889e-3f);
    cp = fmaf(a2, cp, 4.16666666666666667e-2f);
    cp = fmaf(a2, cp, -0.5f);
    const float ca = fmaf(a2, cp, 1.0f);
    c = q == 0 ? ca : q == 1 ? -sa : q == 2 ? -ca : sa;
    s = q == 0 ? sa : q == 1 ? ca : q == 2 ? -sa : -ca;
}

// Cosine-weighted direction about the unit normal n (Malley: uniform disk point lifted to the
// hemisphere) in the orthonormal basis of Duff et al. 2017, normalised.
RT_DEV void cosine_dir(float nx, float ny, float nz, float u0, float u1, float& dx, float& dy,
                       float& dz) {
    const float rad = sqrtf(u0);
    float cs, sn;
    cos_sin_2pi(u1, cs, sn);
    const float lx = rad * cs, ly = rad * sn, lz = sqrtf(fmaxf(0.0f, 1.0f - u0));
    const float sg = copysignf(1.0f, nz);
    const float a = -1.0f / (sg + nz);
    const float b = nx * ny * a;
    const float tx = fmaf(sg * nx * nx, a, 1.0f), ty = sg * b, tz = -sg * nx;
    const float bx = b, by = fmaf(ny * ny, a, sg), bz = -ny;
    const float x = fmaf(lx, tx, fmaf(ly, bx, lz * nx));
    const float y = fmaf(lx, ty, fmaf(ly, by, lz * ny));
    const float z = fmaf(lx, tz, fmaf(ly, bz, lz * nz));
    const float il = 1.0f / sqrtf(fmaf(x, x, fmaf(y, y, z * z)));
    dx = x * il;
    dy = y * il;
    dz = z * il;
}

// Unit geometric normal of a triangle record (e1 x e2), turned to face against direction d.
RT_DEV void facing_normal(const float4 e1, const float4 e2, float dx, float dy, float dz, float& nx,
                          float& ny, float& nz) {
    float x = fmaf(e1.y, e2.z, -(e1.z * e2.y));
    float y = fmaf(e1.z, e2.x, -(e1.x * e2.z));
    float z = fmaf(e1.x, e2.y, -(e1.y * e2.x));
    const float il = 1.0f / sqrtf(fmaf(x, x, fmaf(y, y, z * z)));
    x *= il;
    y *= il;
    z *= il;
    const bool flip = fmaf(x, dx, fmaf(y, dy, z * dz)) > 0.0f;
    nx = flip ? -x : x;
    ny = flip ? -y : y;
    nz = flip ? -z : z;
}

struct PathArgs {
    uint32_t W, H, spp, bounces, key0, key1;
    float albedo, sky, tmin;
    float* samples;        // W*H*spp, sample id = pixel * spp + s
    float4* q[2];          // ray queues: (o.xyz, throughput), (d.xyz, id bits), (skip bits, hit slot
                           // bits, hit t, 0); the hit fields are written by k_paths_bounce
    uint32_t* counts;      // [0] camera rays, [k] rays of depth k (k = 1..bounces)
    uint32_t cap;          // queue capacity (W*H*spp)
};

// The bounce ray leaving the closest hit h of ray r (sample `id` at depth k+1): appended to queue
// `qo` by the lanes with `go` set (compaction: ballot + one atomic per wave).
// The queue entry of the bounce of a hit (go): origin and throughput, direction and sample id, the
// hit's slot (skipped by the bounce's walk).
RT_DEV void bounce_entry(const SceneArgs& sc, const PathArgs& a, const Ray& r, const Hit& h, bool go,
                         uint32_t id, uint32_t depth, float thr, float4& e0, float4& e1, float4& e2) {
    float ox = 0.0f, oy = 0.0f, oz = 0.0f, dx = 0.0f, dy = 0.0f, dz = 0.0f;
    if (go) {
        const float4* tr = sc.tris + (size_t)h.slot * 3;
        float nx, ny, nz;
        facing_normal(tr[1], tr[2], r.dx, r.dy, r.dz, nx, ny, nz);
        const uint32_t pixel = id / a.spp, s = id - pixel * a.spp;
        const uint4 rnd = philox4x32_10(make_uint4(pixel, s, depth, 0u), a.key0, a.key1);
        cosine_dir(nx, ny, nz, unit_float(rnd.x), unit_float(rnd.y), dx, dy, dz);
        ox = fmaf(h.t, r.dx, r.ox);
        oy = fmaf(h.t, r.dy, r.oy);
        oz = fmaf(h.t, r.dz, r.oz);
    }
    e0 = make_float4(ox, oy, oz, thr * a.albedo);
    e1 = make_float4(dx, dy, dz, __int_as_float((int)id));
    e2 = make_float4(__int_as_float(h.slot), 0.0f, 0.0f, 0.0f);
}

// Appends the bounce of each lane with `go` to queue qo (one atomic per wave, on `counter`).
RT_DEV void emit_bounce(const SceneArgs& sc, const PathArgs& a, const Ray& r, const Hit& h, bool go,
                        uint32_t id, uint32_t depth, float thr, float4* qo, uint32_t* counter) {
    float4 e0, e1, e2;
    bounce_entry(sc, a, r, h, go, id, depth, thr, e0, e1, e2);
    const uint32_t slot = wave_append(counter, go);
    if (go) {
        float4* e = qo + (size_t)slot * 3;
        e[0] = e0;
        e[1] = e1;
        e[2] = e2;
    }
}

// Lane ranks among the set bits of a wave mask (the lane's position in an append).
RT_DEV uint32_t mask_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__global__ void k_paths_init(PathArgs a) {
    if (threadIdx.x <= a.bounces) a.counts[threadIdx.x] = threadIdx.x == 0 ? a.W * a.H * a.spp : 0u;
}

// Camera rays, for every sample in turn (sample s of pixel p jittered by the first two Philox
// words of counter (p, s, 0, 0)); the pixel layout per traversal is below.
template <int PM>
__global__ void __launch_bounds__(256) k_paths_primary(SceneArgs sc, rt_camera cam, PathArgs a,
                                                        uint32_t tiles_x, int lane_stack, int budget,
                                                        bool coherent) {
    // PM == RT_TRAVERSAL_QUAD: quad layout, one wave per workgroup (wave j of XCD x takes the 4x4
    // quarter w = j & 3 of 8x8 block (j >> 2) * 8 + x, row-major, at ((w & 1) * 4, (w >> 1) * 4),
    // lanes 4q..4q+3 its pixel (q & 3, q >> 2)); else a 256-thread block covers 16x16 pixels
    constexpr bool QL = PM == RT_TRAVERSAL_QUAD;
    __shared__ int32_t stk[(uses_lds(PM) && !QL) ? kLaneStack * 256 : 1];
    __shared__ float4 pk[QL ? 1 : 4][kPacketLds];
    // QUAD: the bounce entries of up to 4 samples of the wave's 16 pixels, appended with one queue
    // reservation (one same-address atomic per wave and sample took ~2 ms of a framed 4-spp frame)
    __shared__ float4 stage[QL ? 64 * 3 : 1];
    uint32_t staged = 0;
    const uint32_t lane = threadIdx.x & 63;
    uint32_t wave, blk;
    if constexpr (QL) {
        const uint32_t xcd = blockIdx.x & 7u, j = blockIdx.x >> 3;
        wave = j & 3u;
        blk = (j >> 2) * 8u + xcd;
        if (blk >= tiles_x * ((a.H + 7u) / 8u)) return;
    } else {
        wave = threadIdx.x >> 6;
        blk = blockIdx.x;
    }
    const uint32_t col = QL ? (blk % tiles_x) * 8 + (wave & 1) * 4 + ((lane >> 2) & 3)
                            : (blk % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    const uint32_t row = QL ? (blk / tiles_x) * 8 + (wave >> 1) * 4 + (lane >> 4)
                            : (blk / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
    const bool lead = !QL || (lane & 3) == 0;  // the lane that writes for its pixel
    const bool valid = col < a.W && row < a.H;
    const uint32_t pixel = valid ? row * a.W + col : 0u;
    for (uint32_t s = 0; s < a.spp; ++s) {
        const uint4 rnd = philox4x32_10(make_uint4(pixel, s, 0u, 0u), a.key0, a.key1);
        const float fx = (float)col + unit_float(rnd.x), fy = (float)row + unit_float(rnd.y);
        const float nx = fx / (float)a.W * 2.0f - 1.0f, ny = 1.0f - fy / (float)a.H * 2.0f;
        const float x = fmaf(nx, cam.right[0], fmaf(ny, cam.up[0], cam.forward[0]));
        const float y = fmaf(nx, cam.right[1], fmaf(ny, cam.up[1], cam.forward[1]));
        const float z = fmaf(nx, cam.right[2], fmaf(ny, cam.up[2], cam.forward[2]));
        const float il = 1.0f / sqrtf(fmaf(x, x, fmaf(y, y, z * z)));
        Ray r;
        r.ox = cam.eye[0];
        r.oy = cam.eye[1];
        r.oz = cam.eye[2];
        r.dx = x * il;
        r.dy = y * il;
        r.dz = z * il;
        r.tmin = 0.0f;
        r.tmax = INFINITY;
        Hit h;
        if constexpr (QL) {
            // leaf records in pairs: 94 VGPRs and 5 waves per SIMD instead of 111 and 4; the 4-spp
            // frames of the 2M-triangle cluster take 2.5% less time this way (the camera and render
            // kernels gain nothing from pairs on the C2 frame and lose 4% on C1, so they keep 4)
            traverse_quad<false, 2>(sc.nodes4, sc.tris, sc.root, r, valid, -1, h);
        } else {
            trace_primary<PM>(sc, r, valid, h, pk[wave], stk + threadIdx.x, 256, lane_stack, budget,
                              coherent);
        }
        const bool hit = valid && h.prim >= 0;
        if (hit) refine_hit(sc.tris, r, h);
        const uint32_t id = pixel * a.spp + s;
        if (valid && lead) a.samples[id] = hit ? 0.0f : a.sky;
        if (a.bounces > 0) {
            if constexpr (QL) {
                const bool go = hit && lead;
                float4 e0, e1, e2;
                bounce_entry(sc, a, r, h, go, id, 1u, 1.0f, e0, e1, e2);
                const uint64_t mask = __ballot(go);
                if (go) {
                    const uint32_t j = (staged + mask_rank(mask)) * 3u;
                    stage[j] = e0;
                    stage[j + 1] = e1;
                    stage[j + 2] = e2;
                }
                staged += (uint32_t)__popcll(mask);
                if ((s & 3u) == 3u || s + 1u == a.spp) {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    uint32_t b = 0;
                    if (staged && lane == 0) b = atomicAdd(&a.counts[1], staged);
                    b = (uint32_t)__shfl((int)b, 0);
                    for (uint32_t x = lane; x < staged * 3u; x += 64u) a.q[0][(size_t)b * 3 + x] = stage[x];
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    staged = 0;
                }
            } else {
                emit_bounce(sc, a, r, h, hit && lead, id, 1u, 1.0f, a.q[0], &a.counts[1]);
            }
        }
    }
}

// Rays of depth k, queue (k-1)&1 (its count written by the previous kernel on the stream): per-lane
// walks of the 4-wide nodes with refill (refill_walks), the grid's waves sharing the queue's 64-ray
// groups. Each finished ray records the triangle slot it hit (-1: none) and its FP32 t in its own
// queue entry (words 1 and 2 of its third float4); k_paths_emit then shades the queue in order. Keeping the sampling
// and emission out of the walk loop keeps the walk's registers few.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) k_paths_bounce(SceneArgs sc, PathArgs a, uint32_t k,
                                                      int lane_stack) {
    __shared__ int32_t stk[kLaneStack * 64];
    __shared__ float4 pk[kPacketLds];
    __shared__ uint32_t ovf[kRefillList];
    const uint32_t n = a.counts[k];
    float4* qi = a.q[(k - 1) & 1];
    // parked leaves (RT_PARK_PATHS) take ~3 more VGPRs: the waves-per-EU bound keeps the kernel at
    // 128 (4 waves per SIMD) without scratch
    refill_walks<false, RT_TRAVERSAL_LANE, RT_PARK_PATHS, true, false>(
        sc, n, gridDim.x, refill_slot(blockIdx.x, 0, 1, gridDim.x), stk + threadIdx.x, 64, lane_stack, ovf, pk,
        [&](uint32_t i, Ray& r) -> int32_t {
            const float4 e0 = qi[(size_t)i * 3], e1 = qi[(size_t)i * 3 + 1], e2 = qi[(size_t)i * 3 + 2];
            r.ox = e0.x;
            r.oy = e0.y;
            r.oz = e0.z;
            r.dx = e1.x;
            r.dy = e1.y;
            r.dz = e1.z;
            r.tmin = a.tmin;
            r.tmax = INFINITY;
            return __float_as_int(e2.x);
        },
        [&](bool fin, uint32_t i, const Ray&, Hit& h) {
            if (fin) {
                float* e2 = (float*)(qi + (size_t)i * 3 + 2);
                e2[1] = __int_as_float(h.slot);
                e2[2] = h.t;  // the walk's FP32 t, kept where the FP64 refinement declines
            }
        });
}

// The rays of depth k after k_paths_bounce, in queue order: a miss writes its sample; a hit is
// refined in FP64 and appends its bounce to queue k&1 (if k < bounces). All lanes of a wave run
// the same iterations (the emission is wave-wide).
__global__ void __launch_bounds__(256) k_paths_emit(SceneArgs sc, PathArgs a, uint32_t k) {
    const uint32_t n = a.counts[k];
    const float4* qi = a.q[(k - 1) & 1];
    float4* qo = a.q[k & 1];
    const uint32_t stride = gridDim.x * blockDim.x;
    // One queue reservation per wave: the wave first counts the hits among all the entries it will
    // shade (their slot words), reserves that many entries of the next queue with one atomic, and
    // then appends its bounces there in order. Appending per iteration took one same-address
    // atomic per 64 entries, which set the kernel's time; the queue order does not enter any result.
    const bool emits = k < a.bounces;
    uint32_t wbase = 0;
    if (emits) {
        uint32_t cnt = 0;
        for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
            const uint32_t i = base + threadIdx.x;
            const bool hit = i < n && __float_as_int(qi[(size_t)i * 3 + 2].y) >= 0;
            cnt += (uint32_t)__popcll(__ballot(hit));
        }
        if (cnt && (threadIdx.x & 63u) == 0u) wbase = atomicAdd(&a.counts[k + 1], cnt);
        wbase = (uint32_t)__shfl((int)wbase, 0);
    }
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < n;
        Ray r;
        r.ox = r.oy = r.oz = 0.0f;
        r.dx = 1.0f;
        r.dy = r.dz = 0.0f;
        Hit h;
        h.t = INFINITY;
        h.u = h.v = 0.0f;
        h.prim = h.slot = -1;
        float thr = 0.0f;
        uint32_t id = 0;
        if (valid) {
            const float4 e0 = qi[(size_t)i * 3], e1 = qi[(size_t)i * 3 + 1], e2 = qi[(size_t)i * 3 + 2];
            r.ox = e0.x;
            r.oy = e0.y;
            r.oz = e0.z;
            thr = e0.w;
            r.dx = e1.x;
            r.dy = e1.y;
            r.dz = e1.z;
            id = (uint32_t)__float_as_int(e1.w);
            h.slot = __float_as_int(e2.y);
            h.t = e2.z;
        }
        r.tmin = a.tmin;
        r.tmax = INFINITY;
        const bool hit = valid && h.slot >= 0;
        if (valid && !hit) a.samples[id] = thr * a.sky;
        if (hit) refine_hit(sc.tris, r, h);
        if (emits) {
            float4 e0, e1, e2;
            bounce_entry(sc, a, r, h, hit, id, k + 1, thr, e0, e1, e2);
            const uint64_t mask = __ballot(hit);
            if (hit) {
                float4* e = qo + (size_t)(wbase + mask_rank(mask)) * 3;
                e[0] = e0;
                e[1] = e1;
                e[2] = e2;
            }
            wbase += (uint32_t)__popcll(mask);
        }
    }
}

__global__ void __launch_bounds__(256) k_paths_resolve(PathArgs a, float* __restrict__ radiance,
                                                        uint32_t* __restrict__ counts_out) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < (uint64_t)a.W * a.H) {
        float sum = 0.0f;
        for (uint32_t s = 0; s < a.spp; ++s) sum += a.samples[p * a.spp + s];
        radiance[p] = sum / (float)a.spp;
    }
    if (counts_out && p <= a.bounces) counts_out[p] = a.counts[p];
}

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" uint64_t rt_paths_workspace_bytes(uint32_t width, uint32_t height, uint32_t spp) {
    const uint64_t n = (uint64_t)width * height * spp;
    return 256 + ((4 * n + 255) & ~255ull) + 2 * 48 * n;
}

extern "C" int rt_trace_paths(const rt_scene* scene, const rt_camera* cam, const rt_paths_params* p,
                              const rt_paths_out* out, void* workspace, uint64_t ws_bytes,
                              const rt_tuning* tuning, void* stream) {
    int e = check_scene(scene, "rt_trace_paths");
    if (e) return e;
    if (!cam || !p || !out) return set_error(RT_EINVAL, "rt_trace_paths: NULL argument");
    const uint64_t n = (uint64_t)p->width * p->height * p->spp;
    if (p->spp < 1 || p->spp > 1024) return set_error(RT_EINVAL, "rt_trace_paths: spp must be 1..1024");
    if (p->bounces > RT_PATHS_MAX_BOUNCES)
        return set_error(RT_EINVAL, "rt_trace_paths: more than RT_PATHS_MAX_BOUNCES bounces");
    if (n == 0) return RT_OK;
    if (n > 0x7FFFFFFFull) return set_error(RT_ECAPACITY, "rt_trace_paths: width*height*spp >= 2^31");
    if (!out->radiance) return set_error(RT_EINVAL, "rt_trace_paths: radiance output is NULL");
    rt_tuning tu;
    if ((e = resolve_tuning(tuning, tu, "rt_trace_paths"))) return e;
    if (!workspace || ws_bytes < rt_paths_workspace_bytes(p->width, p->height, p->spp))
        return set_error(RT_EINVAL, "rt_trace_paths: workspace too small");
    if ((uintptr_t)workspace & 255)
        return set_error(RT_EINVAL, "rt_trace_paths: workspace must be 256-byte aligned");
    PathArgs a;
    a.W = p->width;
    a.H = p->height;
    a.spp = p->spp;
    a.bounces = p->bounces;
    a.key0 = p->seed;
    a.key1 = p->frame;
    a.albedo = p->albedo;
    a.sky = p->sky;
    a.tmin = p->ray_tmin;
    char* w = (char*)workspace;
    a.counts = (uint32_t*)w;
    a.samples = (float*)(w + 256);
    a.q[0] = (float4*)(w + 256 + ((4 * n + 255) & ~255ull));
    a.q[1] = a.q[0] + 3 * n;
    a.cap = (uint32_t)n;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_paths_init, dim3(1), dim3(64), 0, s, a);
    const SceneArgs sc = scene_args(scene);
    const uint32_t tiles_x = (p->width + 15) / 16, tiles_y = (p->height + 15) / 16;
    const bool coherent = coherent_tiles(*cam, p->width, p->height);
    using K = decltype(&k_paths_primary<0>);
    static const K prim[4] = {k_paths_primary<0>, k_paths_primary<1>, k_paths_primary<2>,
                              k_paths_primary<3>};
    if (tu.primary == RT_TRAVERSAL_QUAD) {
        const uint32_t tx8 = (p->width + 7) / 8, ty8 = (p->height + 7) / 8;
        hipLaunchKernelGGL(k_paths_primary<RT_TRAVERSAL_QUAD>, dim3(32u * ((tx8 * ty8 + 7u) / 8u)), dim3(64), 0, s,
                           sc, *cam, a, tx8, (int)tu.lane_stack, (int)tu.packet_budget, coherent);
    } else {
        hipLaunchKernelGGL(prim[tu.primary], dim3(tiles_x * tiles_y), dim3(256), 0, s, sc, *cam, a,
                           tiles_x, (int)tu.lane_stack, (int)tu.packet_budget, coherent);
    }
    const uint32_t grid = (uint32_t)device_cu_count() * 16u;
    for (uint32_t k = 1; k <= p->bounces; ++k) {
        hipLaunchKernelGGL(k_paths_bounce, dim3(grid), dim3(64), 0, s, sc, a, k, (int)tu.lane_stack);
        hipLaunchKernelGGL(k_paths_emit, dim3(grid), dim3(256), 0, s, sc, a, k);
    }
    const uint64_t px = (uint64_t)p->width * p->height;
    const uint32_t rb = (uint32_t)((std::max<uint64_t>(px, RT_PATHS_MAX_BOUNCES + 1) + 255) / 256);
    hipLaunchKernelGGL(k_paths_resolve, dim3(rb), dim3(256), 0, s, a, out->radiance, out->counts);
    return check_launch("rt_trace_paths");
}
