// rt_common.h — device records, FP32 intersection routines and host error plumbing shared by the
// gfx950 kernels of librtx_hip.so.
//
// Scene records (all 16-byte aligned, read with 128-bit loads):
//   node  (64 B, 4 x float4)  [0] = (b0.lo.x, b0.hi.x, b0.lo.y, b0.hi.y)
//                             [1] = (b1.lo.x, b1.hi.x, b1.lo.y, b1.hi.y)
//                             [2] = (b0.lo.z, b0.hi.z, b1.lo.z, b1.hi.z)
//                             [3] = (ref0, ref1, 0, 0) as int32 bits
//         b0/b1 are the boxes of the two children; ref >= 0 is an internal node, ref < 0 is the
//         leaf of triangle slot ~ref. A traversal step reads one node and tests both children.
//   tri   (48 B, 3 x float4)  (v0.xyz, global primitive id bits), (e1.xyz, 0), (e2.xyz, 0)
//         e1 = v1 - v0, e2 = v2 - v0 in FP32, stored in slot (Morton) order.
//   shade (64 B, 4 x float4)  N0.xyz, N1.xyz, N2.xyz, uv0, uv1, uv2, mesh id bits (slot order)
//
// Arithmetic is written with explicit fmaf and the library is compiled with -ffp-contract=off
// and correctly rounded division/sqrt, so every FP32 result is a fixed function of the inputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/rendertoy_rt.h"

#define RT_DEV __device__ __forceinline__

namespace rt {

constexpr int kWideStack = 128;        // >= 3 * 31: pushes per level x levels of the 4-wide tree
constexpr int kPacketLds = 48;         // float4 per wave: triangles of 4 leaves staged by a node step
constexpr float kTfarScale = 1.0000004f;  // 1 + 2*gamma(3) rounded up (Ize 2013, robust slabs)

// Leaves: a subtree holding at most kLeafMax consecutive slots is one leaf, referenced as
// ~(first | (count - 1) << 29), so traversal refs are < 0 for leaves and node ids otherwise.
constexpr int kLeafMax = 4;
constexpr uint32_t kMaxTris = (1u << 29) - 1;

__host__ __device__ inline int32_t leaf_ref(uint32_t first, uint32_t count) {
    return ~(int32_t)(first | ((count - 1u) << 29));
}
__host__ __device__ inline uint32_t leaf_first(int32_t ref) { return (uint32_t)(~ref) & 0x1FFFFFFFu; }
__host__ __device__ inline uint32_t leaf_count(int32_t ref) { return ((uint32_t)(~ref) >> 29) + 1u; }

struct Ray {
    float ox, oy, oz, dx, dy, dz, tmin, tmax;
};

struct Hit {
    float t, u, v;
    int32_t prim, slot;
};

// A pointer read from device memory (a job table, a texture table) is a generic pointer, and its
// loads and stores compile to flat instructions, which the compiler must wait for as both vector-
// memory and LDS/scalar operations. Every such pointer here addresses global memory: rebuilding it
// as an address-space-1 pointer from its integer value lets the loads and stores through it be
// global ones (a cast from generic and back would fold away).
template <typename T>
RT_DEV T* as_global(T* p) {
    return (T*)(__attribute__((address_space(1))) T*)(uintptr_t)p;
}

RT_DEV float safe_inv(float d) {
    float dd = fabsf(d) < 0x1p-80f ? copysignf(0x1p-80f, d) : d;
    return 1.0f / dd;
}

// Two-sided Moller-Trumbore (Moller & Trumbore 1997) in FP32 over a (v0, e1, e2) record. The
// barycentric tests compare the numerators with the determinant, sign-normalised (u = un/det in
// [0, 1] <=> 0 <= sg*un <= |det| with sg = sign(det); multiplying by +-1 is exact), so a rejection
// costs no division; t, u, v are divided out only for a hit.
RT_DEV bool intersect_tri(const float4 a, const float4 b, const float4 c, const Ray& r, float& t,
                          float& u, float& v) {
    const float px = fmaf(r.dy, c.z, -(r.dz * c.y));
    const float py = fmaf(r.dz, c.x, -(r.dx * c.z));
    const float pz = fmaf(r.dx, c.y, -(r.dy * c.x));
    const float det = fmaf(b.x, px, fmaf(b.y, py, b.z * pz));
    const float sg = det > 0.0f ? 1.0f : -1.0f, ad = det * sg;
    const float sx = r.ox - a.x, sy = r.oy - a.y, sz = r.oz - a.z;
    const float un = fmaf(sx, px, fmaf(sy, py, sz * pz));
    const float us = un * sg;
    if ((det == 0.0f) | (us < 0.0f) | (us > ad)) return false;  // one branch for the three tests
    const float qx = fmaf(sy, b.z, -(sz * b.y));
    const float qy = fmaf(sz, b.x, -(sx * b.z));
    const float qz = fmaf(sx, b.y, -(sy * b.x));
    const float vn = fmaf(r.dx, qx, fmaf(r.dy, qy, r.dz * qz));
    const float vs = vn * sg;
    if ((vs < 0.0f) | (us + vs > ad)) return false;
    const float inv = 1.0f / det;
    t = fmaf(c.x, qx, fmaf(c.y, qy, c.z * qz)) * inv;
    u = un * inv;
    v = vn * inv;
    return true;
}

// FP64 re-evaluation of the winning triangle's (t, u, v): the selection stays FP32 (closest FP32
// Moller-Trumbore t), the reported hit is accurate to FP64 of the stored record. FP32 t degrades
// on sliver triangles far from the origin; this keeps the reported distance within 1e-5 relative.
RT_DEV void refine_hit(const float4* __restrict__ tris, const Ray& r, Hit& h) {
    const float4* tr = tris + (size_t)h.slot * 3;
    const float4 a = tr[0], b = tr[1], c = tr[2];
    const double dx = r.dx, dy = r.dy, dz = r.dz;
    const double px = fma(dy, (double)c.z, -(dz * (double)c.y));
    const double py = fma(dz, (double)c.x, -(dx * (double)c.z));
    const double pz = fma(dx, (double)c.y, -(dy * (double)c.x));
    const double det = fma((double)b.x, px, fma((double)b.y, py, (double)b.z * pz));
    if (det == 0.0) return;
    const double sx = (double)r.ox - (double)a.x, sy = (double)r.oy - (double)a.y,
                 sz = (double)r.oz - (double)a.z;
    const double qx = fma(sy, (double)b.z, -(sz * (double)b.y));
    const double qy = fma(sz, (double)b.x, -(sx * (double)b.z));
    const double qz = fma(sx, (double)b.y, -(sy * (double)b.x));
    const double un = fma(sx, px, fma(sy, py, sz * pz));
    const double vn = fma(dx, qx, fma(dy, qy, dz * qz));
    const double tn = fma((double)c.x, qx, fma((double)c.y, qy, (double)c.z * qz));
    h.t = (float)(tn / det);
    h.u = (float)(un / det);
    h.v = (float)(vn / det);
}

// Tests the triangles of leaf `ref` in slot order against r, updating h (closest hit; ties to the
// smaller primitive id). Every record of the leaf is fetched before the first test (slots past the
// leaf's count re-read its first record), so a leaf costs one memory round trip instead of one per
// triangle. `active` = this lane takes part; `skip_slot` = a slot never reported (a shadow ray's
// own triangle). Returns whether h changed.
RT_DEV bool test_leaf(const float4* __restrict__ tris, int32_t ref, const Ray& r, bool active,
                      int32_t skip_slot, Hit& h) {
    const uint32_t first = leaf_first(ref), count = leaf_count(ref);
    float4 a[kLeafMax], b[kLeafMax], c[kLeafMax];
#pragma unroll
    for (int k = 0; k < kLeafMax; ++k) {
        const float4* tr = tris + (size_t)(first + ((uint32_t)k < count ? (uint32_t)k : 0u)) * 3;
        a[k] = tr[0];
        b[k] = tr[1];
        c[k] = tr[2];
    }
    bool changed = false;
#pragma unroll
    for (int k = 0; k < kLeafMax; ++k) {
        const int32_t slot = (int32_t)(first + (uint32_t)k);
        float t, u, v;
        if ((uint32_t)k < count && active && slot != skip_slot &&
            intersect_tri(a[k], b[k], c[k], r, t, u, v) && t > r.tmin) {
            const int32_t prim = __float_as_int(a[k].w);
            if (t < h.t || (t == h.t && prim < h.prim)) {
                h.t = t;
                h.u = u;
                h.v = v;
                h.prim = prim;
                h.slot = slot;
                changed = true;
            }
        }
    }
    return changed;
}

// Resumable state of a per-lane walk (traverse_bin / traverse_lane4): the ray's slab terms, the
// node it visits next (RT_EMPTY_ROOT: none left), its stack depth (the entries live in the caller's
// LDS stack) and the leaf it has parked for the wave's next leaf round (lane4_steps; RT_EMPTY_ROOT:
// none).
struct LaneWalk {
    float idx, idy, idz, oxx, oyy, ozz;
    int32_t cur;
    int sp;
    int32_t leaf;
};

RT_DEV void walk_start(const Ray& r, int32_t root, LaneWalk& w) {
    w.idx = safe_inv(r.dx);
    w.idy = safe_inv(r.dy);
    w.idz = safe_inv(r.dz);
    w.oxx = r.ox * w.idx;
    w.oyy = r.oy * w.idy;
    w.ozz = r.oz * w.idz;
    w.cur = root;
    w.sp = 0;
    w.leaf = RT_EMPTY_ROOT;
}

// What a bounded run of a per-lane walk ended with.
enum { kWalkOn = 0, kWalkDone = 1, kWalkOverflow = 2 };

// At most `steps` iterations (one node or one leaf each) of the binary walk from state w:
// kWalkDone when the walk has ended (h final), kWalkOverflow when a push would exceed `limit`
// entries (h invalid), else kWalkOn (w and h hold the walk so far).
template <bool ANY>
RT_DEV int bin_steps(const float4* __restrict__ nodes, const float4* __restrict__ tris, const Ray& r,
                     int32_t skip_slot, Hit& h, LaneWalk& w, int32_t* stk, int stride, int limit,
                     int steps) {
    for (int s = 0; s < steps; ++s) {
        if (w.cur >= 0) {
            const float4* nd = nodes + (size_t)w.cur * 4;
            const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
            const float a0 = fmaf(n0.x, w.idx, -w.oxx), a1 = fmaf(n0.y, w.idx, -w.oxx);
            const float a2 = fmaf(n0.z, w.idy, -w.oyy), a3 = fmaf(n0.w, w.idy, -w.oyy);
            const float a4 = fmaf(n2.x, w.idz, -w.ozz), a5 = fmaf(n2.y, w.idz, -w.ozz);
            const float b0 = fmaf(n1.x, w.idx, -w.oxx), b1 = fmaf(n1.y, w.idx, -w.oxx);
            const float b2 = fmaf(n1.z, w.idy, -w.oyy), b3 = fmaf(n1.w, w.idy, -w.oyy);
            const float b4 = fmaf(n2.z, w.idz, -w.ozz), b5 = fmaf(n2.w, w.idz, -w.ozz);
            const float tn0 = fmaxf(fmaxf(fminf(a0, a1), fminf(a2, a3)), fmaxf(fminf(a4, a5), r.tmin));
            const float tf0 = fminf(fminf(fmaxf(a0, a1), fmaxf(a2, a3)), fmaxf(a4, a5)) * kTfarScale;
            const float tn1 = fmaxf(fmaxf(fminf(b0, b1), fminf(b2, b3)), fmaxf(fminf(b4, b5), r.tmin));
            const float tf1 = fminf(fminf(fmaxf(b0, b1), fmaxf(b2, b3)), fmaxf(b4, b5)) * kTfarScale;
            const bool hit0 = tn0 <= fminf(tf0, h.t);
            const bool hit1 = tn1 <= fminf(tf1, h.t);
            const int32_t r0 = __float_as_int(n3.x), r1 = __float_as_int(n3.y);
            if (hit0 && hit1) {
                if (w.sp >= limit) return kWalkOverflow;
                const bool swap = tn1 < tn0;
                stk[(w.sp++) * stride] = swap ? r0 : r1;
                w.cur = swap ? r1 : r0;
                continue;
            }
            if (hit0 | hit1) {
                w.cur = hit0 ? r0 : r1;
                continue;
            }
        } else {
            if (test_leaf(tris, w.cur, r, true, skip_slot, h) && ANY) return kWalkDone;
        }
        if (w.sp == 0) return kWalkDone;
        w.cur = stk[(--w.sp) * stride];
    }
    return kWalkOn;
}

// Closest-hit (ANY=false) or any-hit (ANY=true) per-lane traversal of the two-child node layout,
// nearer child first, pending children on a per-lane stack in LDS: entry k of this lane lives at
// stk[k * stride] (stride = threads per block, so the lanes of a wave touch consecutive words at
// equal depth and every access is one LDS instruction with no bank conflict). `skip_slot` is never
// tested (a shadow ray's own triangle). Returns false if the ray needed more than `limit` pending
// entries: the hit is then invalid and the caller traces the ray again with a traversal whose stack
// always suffices. On a true return h.prim < 0 means no hit in (tmin, tmax).
template <bool ANY>
RT_DEV bool traverse_bin(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                         int32_t root, const Ray& r, int32_t skip_slot, Hit& h, int32_t* stk,
                         int stride, int limit) {
    h.t = r.tmax;
    h.u = 0.0f;
    h.v = 0.0f;
    h.prim = -1;
    h.slot = -1;
    if (root == RT_EMPTY_ROOT) return true;
    LaneWalk w;
    walk_start(r, root, w);
    return bin_steps<ANY>(nodes, tris, r, skip_slot, h, w, stk, stride, limit, 0x7FFFFFFF) != kWalkOverflow;
}

// Wave-coherent ("packet") traversal of the 4-wide nodes for coherent rays (camera tiles, shadow
// rays toward one light from a tile's hit points). The 64 lanes walk one node sequence together:
//  - the current node, the stack pointer and the 64-bit mask of lanes that passed the current
//    node's box are wave-uniform, so node records are fetched with scalar loads;
//  - each lane tests only boxes and triangles of nodes it entered itself (the mask), with its own
//    t culling, so a lane's result is that of its own ray;
//  - the leaf children a node step finds hit are tested in that same step: one vector load stages
//    the triangles of all of them (at most 4 leaves x 4 triangles x 3 float4 = 48 lanes) in the
//    wave's 48-float4 LDS area `lds`, and every lane reads them from there. Only internal nodes are
//    pushed, and the record of the node visited next (the nearest internal child, else the stack
//    top) is requested before the leaf tests, so its fetch overlaps them: a step costs one memory
//    round trip;
//  - internal children are visited in the order of the entry distances of the lowest lane of the
//    mask (children that lane misses after those it hits, ties by child index);
//  - the stack is a pair of VGPRs per field, lane k holding entries k and 64+k (push = per-lane
//    select, pop = v_readlane): at most 3 pushes on each of at most 31 levels.
// `budget` > 0 caps the node steps of the wave: the walk stops after that many and returns false
// (h then holds the best hit found so far, a valid bound for a restart); otherwise it runs to the
// end and returns true. Must be called by all 64 lanes of the wave with uniform control flow (no
// lane may have exited).
template <bool ANY>
RT_DEV bool traverse_packet(const float4* __restrict__ nodes4, const float4* __restrict__ tris,
                            int32_t root, const Ray& r, bool want, int32_t skip_slot, Hit& h,
                            float4* lds, int budget = 0) {
    h.t = r.tmax;
    h.u = 0.0f;
    h.v = 0.0f;
    h.prim = -1;
    h.slot = -1;
    uint64_t m = __ballot(want);
    if (root == RT_EMPTY_ROOT || m == 0) return true;
    const uint32_t lane = __lane_id();
    if (root < 0) {  // the whole tree is one leaf
        test_leaf(tris, root, r, ((m >> lane) & 1u) != 0, skip_slot, h);
        return true;
    }
    uint64_t alive = m;
    const float idx = safe_inv(r.dx), idy = safe_inv(r.dy), idz = safe_inv(r.dz);
    const float oxx = r.ox * idx, oyy = r.oy * idy, ozz = r.oz * idz;
    int32_t stk_node0 = 0, stk_node1 = 0;
    uint32_t stk_mlo0 = 0, stk_mlo1 = 0, stk_mhi0 = 0, stk_mhi1 = 0;
    int32_t sp = 0;
    // pops entries until one has lanes left (after `alive` for any-hit); false if the stack is empty
    auto pop = [&](int32_t& node, uint64_t& mask) -> bool {
        for (;;) {
            if (sp == 0) return false;
            --sp;
            const bool hi = sp >= 64;
            const int32_t lsp = hi ? sp - 64 : sp;
            node = __builtin_amdgcn_readlane(hi ? stk_node1 : stk_node0, lsp);
            const uint32_t mlo = (uint32_t)__builtin_amdgcn_readlane((int)(hi ? stk_mlo1 : stk_mlo0), lsp);
            const uint32_t mhi = (uint32_t)__builtin_amdgcn_readlane((int)(hi ? stk_mhi1 : stk_mhi0), lsp);
            mask = ((uint64_t)mhi << 32) | mlo;
            if (ANY) mask &= alive;
            if (mask != 0) return true;
        }
    };
    const float4* nd = nodes4 + (size_t)root * 8;
    float4 lx = nd[0], hx = nd[1], ly = nd[2], hy = nd[3], lz = nd[4], hz = nd[5], rf = nd[6];
    int steps = 0;
    for (;;) {
        if (ANY) {
            m &= alive;
            if (m == 0) {  // every lane of this entry found its occluder meanwhile
                int32_t node;
                if (!pop(node, m)) return true;
                nd = nodes4 + (size_t)node * 8;
                lx = nd[0]; hx = nd[1]; ly = nd[2]; hy = nd[3]; lz = nd[4]; hz = nd[5]; rf = nd[6];
            }
        }
        if (budget > 0 && ++steps > budget) return false;
        const bool mine = (m >> lane) & 1u;
        const float clx[4] = {lx.x, lx.y, lx.z, lx.w}, chx[4] = {hx.x, hx.y, hx.z, hx.w};
        const float cly[4] = {ly.x, ly.y, ly.z, ly.w}, chy[4] = {hy.x, hy.y, hy.z, hy.w};
        const float clz[4] = {lz.x, lz.y, lz.z, lz.w}, chz[4] = {hz.x, hz.y, hz.z, hz.w};
        const int32_t ref[4] = {__float_as_int(rf.x), __float_as_int(rf.y), __float_as_int(rf.z),
                                __float_as_int(rf.w)};
        struct Child {
            float key;
            int32_t k, ref;
            uint64_t mask;
        };
        Child c[4];
        uint64_t lmask[4];
        const int rep = __ffsll((unsigned long long)m) - 1;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float x0 = fmaf(clx[k], idx, -oxx), x1 = fmaf(chx[k], idx, -oxx);
            const float y0 = fmaf(cly[k], idy, -oyy), y1 = fmaf(chy[k], idy, -oyy);
            const float z0 = fmaf(clz[k], idz, -ozz), z1 = fmaf(chz[k], idz, -ozz);
            const float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), r.tmin));
            const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1)) * kTfarScale;
            const bool hk = mine & (ref[k] != RT_EMPTY_ROOT) & (tn <= fminf(tf, h.t));  // no short circuit
            const uint64_t mk = __ballot(hk);
            const bool leaf = ref[k] < 0;
            lmask[k] = leaf ? mk : 0;
            c[k].mask = leaf ? 0 : mk;
            c[k].key = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(hk ? tn : INFINITY), rep));
            c[k].k = k;
            c[k].ref = ref[k];
        }
        // internal children by (key, k): 4-element sorting network over whole records
        auto cswap = [](Child& a, Child& b) {
            if (b.key < a.key || (b.key == a.key && b.k < a.k)) {
                const Child t = a;
                a = b;
                b = t;
            }
        };
        cswap(c[0], c[1]);
        cswap(c[2], c[3]);
        cswap(c[0], c[2]);
        cswap(c[1], c[3]);
        cswap(c[1], c[2]);
        int32_t next = -1;
        uint64_t next_m = 0;
#pragma unroll
        for (int q = 3; q >= 0; --q) {  // push the farther ones, the nearest is visited next
            if (c[q].mask == 0) continue;
            if (next_m != 0 && sp < kWideStack) {
                const bool at0 = (int32_t)lane == sp, at1 = (int32_t)lane + 64 == sp;
                stk_node0 = at0 ? next : stk_node0;
                stk_mlo0 = at0 ? (uint32_t)next_m : stk_mlo0;
                stk_mhi0 = at0 ? (uint32_t)(next_m >> 32) : stk_mhi0;
                stk_node1 = at1 ? next : stk_node1;
                stk_mlo1 = at1 ? (uint32_t)next_m : stk_mlo1;
                stk_mhi1 = at1 ? (uint32_t)(next_m >> 32) : stk_mhi1;
                ++sp;
            }
            next = c[q].ref;
            next_m = c[q].mask;
        }
        bool more = next_m != 0;
        if (!more) more = pop(next, next_m);
        // request the next node's record now: its fetch overlaps the leaf tests below
        if (more) {
            nd = nodes4 + (size_t)next * 8;
            lx = nd[0]; hx = nd[1]; ly = nd[2]; hy = nd[3]; lz = nd[4]; hz = nd[5]; rf = nd[6];
        }
        if ((lmask[0] | lmask[1] | lmask[2] | lmask[3]) != 0) {
            // stage: lane j < 48 fetches float4 (j % 12) of leaf child j / 12
            const uint32_t kj = lane / 12u, part = lane - kj * 12u;
            const uint64_t mj = kj == 0 ? lmask[0] : kj == 1 ? lmask[1] : kj == 2 ? lmask[2] : lmask[3];
            const int32_t rj = kj == 0 ? ref[0] : kj == 1 ? ref[1] : kj == 2 ? ref[2] : ref[3];
            if (lane < 48u && mj != 0) {
                const uint32_t t = part / 3u, comp = part - t * 3u, cnt = leaf_count(rj);
                const uint32_t slot = leaf_first(rj) + (t < cnt ? t : 0u);
                lds[lane] = tris[(size_t)slot * 3 + comp];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (lmask[k] == 0) continue;
                const bool act = (lmask[k] >> lane) & 1u;
                const uint32_t first = leaf_first(ref[k]), cnt = leaf_count(ref[k]);
#pragma unroll
                for (uint32_t t = 0; t < (uint32_t)kLeafMax; ++t) {
                    if (t >= cnt) break;
                    const float4 ta = lds[k * 12 + t * 3], tb = lds[k * 12 + t * 3 + 1],
                                 tc = lds[k * 12 + t * 3 + 2];
                    const int32_t slot = (int32_t)(first + t);
                    float tt, u, v;
                    if (act && slot != skip_slot && intersect_tri(ta, tb, tc, r, tt, u, v) && tt > r.tmin) {
                        const int32_t prim = __float_as_int(ta.w);
                        if (tt < h.t || (tt == h.t && prim < h.prim)) {
                            h.t = tt;
                            h.u = u;
                            h.v = v;
                            h.prim = prim;
                            h.slot = slot;
                        }
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (ANY) {
                alive &= ~__ballot(h.prim >= 0);
                if (alive == 0) return true;
            }
        }
        if (!more) return true;
        m = next_m;
    }
}

// ---- quad-lane packets: 16 rays per wave, 4 lanes per ray ----
// Lane 4q+c holds ray q (the 4 lanes of a quad hold the same ray and the same `want`); at each node
// lane c tests child c's box and at each leaf triangle c, so a step costs one box test and at most
// four triangle tests per lane. Each lane keeps its own closest candidate; boxes are culled with the
// quad's best t (a DPP minimum after each leaf step), and the quad's closest hit is combined across
// its 4 lanes with DPP quad permutations once, at the end. The visiting order is the packet
// traversal's (traverse_packet) over the 16 rays: masks hold one bit per ray (lane 4q), internal
// children by the entry distances of the lowest active ray, hit leaf children tested in the node
// step, the next node fetched before the leaf tests. Results are replicated in the 4 lanes of a quad.
constexpr uint64_t kQuadBits = 0x1111111111111111ull;  // lane 4q of every quad

RT_DEV int dpp_xor1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false); }  // [1,0,3,2]
RT_DEV int dpp_xor2(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false); }  // [2,3,0,1]

// closest of the candidates (t, prim) of the 4 lanes of each quad, ties to the smaller prim;
// prim < 0 marks no candidate (t = +inf). Called with all lanes active.
RT_DEV void quad_closest(float& t, float& u, float& v, int32_t& prim, int32_t& slot) {
#pragma unroll
    for (int step = 0; step < 2; ++step) {
        const float ot = __int_as_float(step ? dpp_xor2(__float_as_int(t)) : dpp_xor1(__float_as_int(t)));
        const float ou = __int_as_float(step ? dpp_xor2(__float_as_int(u)) : dpp_xor1(__float_as_int(u)));
        const float ov = __int_as_float(step ? dpp_xor2(__float_as_int(v)) : dpp_xor1(__float_as_int(v)));
        const int32_t op = step ? dpp_xor2(prim) : dpp_xor1(prim);
        const int32_t os = step ? dpp_xor2(slot) : dpp_xor1(slot);
        const bool take = op >= 0 && (prim < 0 || ot < t || (ot == t && op < prim));
        t = take ? ot : t;
        u = take ? ou : u;
        v = take ? ov : v;
        prim = take ? op : prim;
        slot = take ? os : slot;
    }
}

// minimum of t over the 4 lanes of each quad (all lanes active)
RT_DEV float quad_min(float t) {
    t = fminf(t, __int_as_float(dpp_xor1(__float_as_int(t))));
    return fminf(t, __int_as_float(dpp_xor2(__float_as_int(t))));
}

// bit 4q set if any lane of quad q is set in the ballot b
RT_DEV uint64_t quad_any(uint64_t b) { return (b | (b >> 1) | (b >> 2) | (b >> 3)) & kQuadBits; }

// Stack entries as the quad walk keeps them: a node and its 16 ray bits packed in 32 bits (ray q
// at bit 4q for q < 8, at bit 4(q-8)+1 for q >= 8), entry k in lane k (k < 64) or lane k - 64.
RT_DEV uint32_t quad_mask_pack(uint64_t mask) { return (uint32_t)mask | ((uint32_t)(mask >> 32) << 1); }
RT_DEV uint64_t quad_mask_unpack(uint32_t cm) {
    return (uint64_t)(cm & 0x11111111u) | ((uint64_t)((cm >> 1) & 0x11111111u) << 32);
}

// The traversal proper: on return each lane's h is its own closest candidate (see traverse_quad).
// The walk starts at node `start` with the rays of `m` (bit 4q = ray q) and the candidates already
// in h (all four lanes of a quad hold the same one: a bound found elsewhere, or none). Called by
// all 64 lanes with uniform control flow.
template <bool ANY, int LG = 4>
RT_DEV void traverse_quad_core(const float4* __restrict__ nodes4, const float4* __restrict__ tris,
                               int32_t start, uint64_t m, const Ray& r, int32_t skip_slot, Hit& h) {
    const uint32_t lane = __lane_id(), c = lane & 3u, qb = lane & ~3u;
    uint64_t alive = m;
    const float idx = safe_inv(r.dx), idy = safe_inv(r.dy), idz = safe_inv(r.dz);
    const float oxx = r.ox * idx, oyy = r.oy * idy, ozz = r.oz * idz;
    float qt = quad_min(h.t);  // the quad's best t so far (culling); h holds each lane's own best
    // tests triangle c of each leaf in `lref` (ray in the leaf's mask), keeps this lane's closest
    // candidate in h and refreshes the quad's best t. The records of LG leaves are fetched before
    // their tests: LG = 4 takes one memory round trip per step (48 VGPRs of records; the camera
    // kernel at 90 VGPRs, 5 waves per SIMD), LG = 2 at most two (74 VGPRs, 6 waves per SIMD).
    auto leaves = [&](uint32_t lset, const int32_t* lref, const uint64_t* lmsk) {
        float bt = INFINITY, bu = 0.0f, bv = 0.0f;
        int32_t bp = -1, bs = -1;
        float4 ta[4], tb[4], tc[4];
#pragma unroll
        for (int g = 0; g < 4; g += LG) {
#pragma unroll
        for (int k = g; k < g + LG; ++k) {
            if (!((lset >> k) & 1u)) continue;
            const uint32_t first = leaf_first(lref[k]), cnt = leaf_count(lref[k]);
            const float4* tr = tris + (size_t)(first + (c < cnt ? c : 0u)) * 3;
            ta[k] = tr[0];
            tb[k] = tr[1];
            tc[k] = tr[2];
        }
        if (LG < 4) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = g; k < g + LG; ++k) {
            if (!((lset >> k) & 1u)) continue;
            const uint32_t first = leaf_first(lref[k]), cnt = leaf_count(lref[k]);
            const int32_t slot = (int32_t)(first + c);
            float tt, uu, vv;
            if (c < cnt && ((lmsk[k] >> qb) & 1u) && slot != skip_slot &&
                intersect_tri(ta[k], tb[k], tc[k], r, tt, uu, vv) && tt > r.tmin) {
                const int32_t prim = __float_as_int(ta[k].w);
                if (bp < 0 || tt < bt || (tt == bt && prim < bp)) {
                    bt = tt;
                    bu = uu;
                    bv = vv;
                    bp = prim;
                    bs = slot;
                }
            }
        }
        }
        if (bp >= 0 && (bt < h.t || (bt == h.t && bp < h.prim))) {
            h.t = bt;
            h.u = bu;
            h.v = bv;
            h.prim = bp;
            h.slot = bs;
        }
        qt = quad_min(h.t);
    };
    if (start < 0) {  // the whole tree is one leaf
        const int32_t lref[4] = {start, 0, 0, 0};
        const uint64_t lmsk[4] = {m, 0, 0, 0};
        leaves(1u, lref, lmsk);
        return;
    }
    int32_t stk_node0 = 0, stk_node1 = 0;
    uint32_t stk_m0 = 0, stk_m1 = 0;  // ray masks, bit 4q -> bit 4q (q < 8) / 4(q-8)+1
    int32_t sp = 0;
    auto pop = [&](int32_t& node, uint64_t& mask) -> bool {
        for (;;) {
            if (sp == 0) return false;
            --sp;
            const bool hi = sp >= 64;
            const int32_t lsp = hi ? sp - 64 : sp;
            node = __builtin_amdgcn_readlane(hi ? stk_node1 : stk_node0, lsp);
            const uint32_t cm = (uint32_t)__builtin_amdgcn_readlane((int)(hi ? stk_m1 : stk_m0), lsp);
            mask = quad_mask_unpack(cm);
            if (ANY) mask &= alive;
            if (mask != 0) return true;
        }
    };
    // lane c's child of the node being visited: its box and ref (fields c of the SoA record)
    const float* nf = (const float*)(nodes4 + (size_t)start * 8);
    float clx = nf[c], chx = nf[4 + c], cly = nf[8 + c], chy = nf[12 + c], clz = nf[16 + c],
          chz = nf[20 + c];
    int32_t cref = __float_as_int(nf[24 + c]);
    for (;;) {
        if (ANY) {
            m &= alive;
            if (m == 0) {
                int32_t node;
                if (!pop(node, m)) return;
                nf = (const float*)(nodes4 + (size_t)node * 8);
                clx = nf[c]; chx = nf[4 + c]; cly = nf[8 + c]; chy = nf[12 + c]; clz = nf[16 + c];
                chz = nf[20 + c]; cref = __float_as_int(nf[24 + c]);
            }
        }
        const bool mine = (m >> qb) & 1u;
        const float x0 = fmaf(clx, idx, -oxx), x1 = fmaf(chx, idx, -oxx);
        const float y0 = fmaf(cly, idy, -oyy), y1 = fmaf(chy, idy, -oyy);
        const float z0 = fmaf(clz, idz, -ozz), z1 = fmaf(chz, idz, -ozz);
        const float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), r.tmin));
        const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1)) * kTfarScale;
        const bool hk = mine & (cref != RT_EMPTY_ROOT) & (tn <= fminf(tf, qt));  // no short circuit
        const uint64_t mk = __ballot(hk);
        // children some ray entered (bit c of the OR of the 16 quads' nibbles): internal ones (live)
        // and leaves (tested below, in this step)
        uint64_t fold = mk | (mk >> 32);
        fold |= fold >> 16;
        fold |= fold >> 8;
        fold |= fold >> 4;
        const uint32_t lbits = (uint32_t)__ballot(cref < 0);  // bit c: child c is a leaf
        const uint32_t live = (uint32_t)fold & ~lbits & 0xFu, leafs = (uint32_t)fold & lbits & 0xFu;
        int32_t lref[4];
        uint64_t lmsk[4];
        if (leafs != 0) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if ((leafs >> k) & 1u) {
                    lref[k] = __builtin_amdgcn_readlane(cref, k);
                    lmsk[k] = (mk >> k) & kQuadBits;
                }
            }
        }
        auto push = [&](int32_t node, uint64_t mask) {
            if (sp < kWideStack) {
                const bool at0 = (int32_t)lane == sp, at1 = (int32_t)lane + 64 == sp;
                const uint32_t cm = quad_mask_pack(mask);
                stk_node0 = at0 ? node : stk_node0;
                stk_m0 = at0 ? cm : stk_m0;
                stk_node1 = at1 ? node : stk_node1;
                stk_m1 = at1 ? cm : stk_m1;
                ++sp;
            }
        };
        // Internal children in the order of (entry distance of the lowest active ray, k): the
        // nearest is visited next, the others are pushed farthest first. One or two live children
        // (the common case) take a direct path; three or four a 5-comparator network.
        int32_t next = -1;
        uint64_t next_m = 0;
        const int rep = __ffsll((unsigned long long)m) - 1;  // lane 4q of the lowest active ray
        const float keyv = hk ? tn : INFINITY;
        const int nlive = __builtin_popcount(live);
        if (nlive == 1) {
            const int k = __builtin_ctz(live);
            next = __builtin_amdgcn_readlane(cref, k);
            next_m = (mk >> k) & kQuadBits;
        } else if (nlive == 2) {
            const int k0 = __builtin_ctz(live), k1 = 31 - __builtin_clz(live);  // k0 < k1
            const float key0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(keyv), rep + k0));
            const float key1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(keyv), rep + k1));
            const bool swap = key1 < key0;
            const int kn = swap ? k1 : k0, kf = swap ? k0 : k1;
            push(__builtin_amdgcn_readlane(cref, kf), (mk >> kf) & kQuadBits);
            next = __builtin_amdgcn_readlane(cref, kn);
            next_m = (mk >> kn) & kQuadBits;
        } else if (nlive > 2) {
            struct Child {
                float key;
                int32_t k, ref;
                uint64_t mask;
            };
            Child ch[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                ch[k].key = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(keyv), rep + k));
                ch[k].k = k;
                ch[k].ref = __builtin_amdgcn_readlane(cref, k);
                ch[k].mask = (live >> k) & 1u ? (mk >> k) & kQuadBits : 0;
            }
            auto cswap = [](Child& a, Child& b) {
                if (b.key < a.key || (b.key == a.key && b.k < a.k)) {
                    const Child t = a;
                    a = b;
                    b = t;
                }
            };
            cswap(ch[0], ch[1]);
            cswap(ch[2], ch[3]);
            cswap(ch[0], ch[2]);
            cswap(ch[1], ch[3]);
            cswap(ch[1], ch[2]);
#pragma unroll
            for (int q = 3; q >= 0; --q) {
                if (ch[q].mask == 0) continue;
                if (next_m != 0) push(next, next_m);
                next = ch[q].ref;
                next_m = ch[q].mask;
            }
        }
        bool more = next_m != 0;
        if (!more) more = pop(next, next_m);
        if (more) {  // the next node's fetch overlaps the leaf tests
            nf = (const float*)(nodes4 + (size_t)next * 8);
            clx = nf[c]; chx = nf[4 + c]; cly = nf[8 + c]; chy = nf[12 + c]; clz = nf[16 + c];
            chz = nf[20 + c]; cref = __float_as_int(nf[24 + c]);
        }
        if (leafs != 0) {
            leaves(leafs, lref, lmsk);
            if (ANY) {
                alive &= ~quad_any(__ballot(h.prim >= 0));
                if (alive == 0) return;
            }
        }
        if (!more) return;
        m = next_m;
    }
}

// The unbudgeted walk of the 16 rays of a wave from the root (lane 4q+c holds ray q; `want` is the
// same in the 4 lanes of a quad).
template <bool ANY, int LG = 4>
RT_DEV void traverse_quad_lanes(const float4* __restrict__ nodes4, const float4* __restrict__ tris,
                                int32_t root, const Ray& r, bool want, int32_t skip_slot, Hit& h) {
    h.t = r.tmax;
    h.u = 0.0f;
    h.v = 0.0f;
    h.prim = -1;
    h.slot = -1;
    const uint64_t m = __ballot(want) & kQuadBits;
    if (root == RT_EMPTY_ROOT || m == 0) return;
    traverse_quad_core<ANY, LG>(nodes4, tris, root, m, r, skip_slot, h);
}

// Quad-lane traversal of the 16 rays of a wave (lane 4q+c holds ray q); h = the ray's closest hit
// (ANY: a hit) in all 4 lanes of its quad. Called by all 64 lanes.
template <bool ANY, int LG = 4>
RT_DEV void traverse_quad(const float4* __restrict__ nodes4, const float4* __restrict__ tris,
                          int32_t root, const Ray& r, bool want, int32_t skip_slot, Hit& h) {
    traverse_quad_lanes<ANY, LG>(nodes4, tris, root, r, want, skip_slot, h);
    quad_closest(h.t, h.u, h.v, h.prim, h.slot);
}

constexpr int kLaneStack = RT_LANE_STACK_MAX;

// Leaf rounds of the per-lane walk (lane4_steps' PARK): the lanes of a wave test their parked
// leaves together once at least PARK quarters of the walking lanes hold one, or when no lane has a
// node left to visit. PARK = 0: no parking (each iteration visits one node or tests one leaf per
// lane). The defaults per caller:
#ifndef RT_PARK_RENDER
#define RT_PARK_RENDER 2  // render and batch kernels (primary and shadow walks)
#endif
#ifndef RT_PARK_QUERY
#define RT_PARK_QUERY 2   // rt_trace_closest
#endif
#ifndef RT_PARK_QUERY_ANY
#define RT_PARK_QUERY_ANY 2  // rt_trace_any
#endif
#ifndef RT_PARK_PATHS
#define RT_PARK_PATHS 2   // rt_trace_paths' bounce rays
#endif

// The children of 4-wide node `cur` whose boxes ray r (slab terms in w) enters no later than `best`:
// returns how many, with their refs in c[0..n-1] nearest entry first (equal entry distances in any
// order: the visiting order decides only how soon the best hit shrinks, never which hit is found).
// NF: near and far planes chosen by the ray's direction signs instead of min/max. Both planes go
// through the same fma and rounding is monotone in the plane, so min(fma(lo), fma(hi)) is fma(lo)
// when the inverse direction is positive and fma(hi) when it is negative: loading the planes in that
// order gives the same entry and exit values without the min/max, at the cost of one address per
// plane. The bounce walks of rt_trace_paths use it (7% faster there); the other per-lane walks keep
// the min/max form (the batch refill kernel is 5-8% slower with it).
template <bool NF = false>
RT_DEV int lane4_children(const float4* __restrict__ nodes4, int32_t cur, const Ray& r, const LaneWalk& w,
                          float best, int32_t c[4]) {
    const float4* nd = nodes4 + (size_t)cur * 8;
    float cnx[4], cfx[4], cny[4], cfy[4], cnz[4], cfz[4];
    if constexpr (NF) {
        const uint32_t sx = __float_as_uint(w.idx) >> 31, sy = __float_as_uint(w.idy) >> 31,
                       sz = __float_as_uint(w.idz) >> 31;
        const float4 nx = nd[sx], fx = nd[1u - sx], ny = nd[2u + sy], fy = nd[3u - sy], nz = nd[4u + sz],
                     fz = nd[5u - sz];
        cnx[0] = nx.x; cnx[1] = nx.y; cnx[2] = nx.z; cnx[3] = nx.w;
        cfx[0] = fx.x; cfx[1] = fx.y; cfx[2] = fx.z; cfx[3] = fx.w;
        cny[0] = ny.x; cny[1] = ny.y; cny[2] = ny.z; cny[3] = ny.w;
        cfy[0] = fy.x; cfy[1] = fy.y; cfy[2] = fy.z; cfy[3] = fy.w;
        cnz[0] = nz.x; cnz[1] = nz.y; cnz[2] = nz.z; cnz[3] = nz.w;
        cfz[0] = fz.x; cfz[1] = fz.y; cfz[2] = fz.z; cfz[3] = fz.w;
    } else {
        const float4 lx = nd[0], hx = nd[1], ly = nd[2], hy = nd[3], lz = nd[4], hz = nd[5];
        cnx[0] = lx.x; cnx[1] = lx.y; cnx[2] = lx.z; cnx[3] = lx.w;
        cfx[0] = hx.x; cfx[1] = hx.y; cfx[2] = hx.z; cfx[3] = hx.w;
        cny[0] = ly.x; cny[1] = ly.y; cny[2] = ly.z; cny[3] = ly.w;
        cfy[0] = hy.x; cfy[1] = hy.y; cfy[2] = hy.z; cfy[3] = hy.w;
        cnz[0] = lz.x; cnz[1] = lz.y; cnz[2] = lz.z; cnz[3] = lz.w;
        cfz[0] = hz.x; cfz[1] = hz.y; cfz[2] = hz.z; cfz[3] = hz.w;
    }
    const float4 rf = nd[6];
    int32_t ref[4] = {__float_as_int(rf.x), __float_as_int(rf.y), __float_as_int(rf.z), __float_as_int(rf.w)};
    float key[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float x0 = fmaf(cnx[k], w.idx, -w.oxx), x1 = fmaf(cfx[k], w.idx, -w.oxx);
        const float y0 = fmaf(cny[k], w.idy, -w.oyy), y1 = fmaf(cfy[k], w.idy, -w.oyy);
        const float z0 = fmaf(cnz[k], w.idz, -w.ozz), z1 = fmaf(cfz[k], w.idz, -w.ozz);
        float tn, tf;
        if constexpr (NF) {
            tn = fmaxf(fmaxf(x0, y0), fmaxf(z0, r.tmin));
            tf = fminf(fminf(x1, y1), z1) * kTfarScale;
        } else {
            tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), r.tmin));
            tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1)) * kTfarScale;
        }
        const bool hk = (ref[k] != RT_EMPTY_ROOT) & (tn <= fminf(tf, best));
        key[k] = hk ? tn : INFINITY;
    }
    // 4-element sorting network over (key, ref): missed children (key +inf) sort last
#define RT_CSWAP(a, b)                                              \
    {                                                               \
        const bool sw = key[b] < key[a];                            \
        const float ka = key[a], kb = key[b];                       \
        const int32_t ra = ref[a], rb = ref[b];                     \
        key[a] = sw ? kb : ka;                                      \
        key[b] = sw ? ka : kb;                                      \
        ref[a] = sw ? rb : ra;                                      \
        ref[b] = sw ? ra : rb;                                      \
    }
    RT_CSWAP(0, 1)
    RT_CSWAP(2, 3)
    RT_CSWAP(0, 2)
    RT_CSWAP(1, 3)
    RT_CSWAP(1, 2)
#undef RT_CSWAP
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] = ref[k];
    return (key[0] < INFINITY) + (key[1] < INFINITY) + (key[2] < INFINITY) + (key[3] < INFINITY);
}

// At most `steps` iterations of the 4-wide walk from state w (returns as bin_steps). `any_rt`: the
// walk ends at the first hit, as with ANY (a per-lane choice).
// PARK > 0: lanes of a wave sit at inner nodes and at leaves in different iterations, and a wave
// issues the node code and the leaf code (its costlier part) whenever any lane needs either. So a
// lane that reaches a leaf parks it and goes on with its stack, and the leaf code runs once per leaf
// round for all parked leaves (speculative traversal, Aila & Laine 2009). The closest hit does not
// depend on the visiting order (a node is skipped only past the best hit so far, which a later leaf
// can only lower), so the hits are those of the one-node-or-leaf walk. Parking costs ~3 VGPRs.
template <bool ANY, int PARK = RT_PARK_RENDER, bool NF = false>
RT_DEV int lane4_steps(const float4* __restrict__ nodes4, const float4* __restrict__ tris, const Ray& r,
                       int32_t skip_slot, Hit& h, LaneWalk& w, int32_t* stk, int stride, int limit,
                       int steps, bool any_rt = false) {
    static_assert(PARK >= 0 && PARK <= 4, "quarters of the walking lanes");
    for (int s = 0; s < steps; ++s) {
        if constexpr (PARK > 0) {
            if (w.cur >= 0 && w.cur != RT_EMPTY_ROOT) {
                int32_t c[4];
                const int nhit = lane4_children<NF>(nodes4, w.cur, r, w, h.t, c);
                if (nhit > 0) {
                    if (w.sp + nhit - 1 > limit) return kWalkOverflow;
                    // push the farther children (farthest first), go on with the nearest
                    if (nhit >= 4) stk[(w.sp++) * stride] = c[3];
                    if (nhit >= 3) stk[(w.sp++) * stride] = c[2];
                    if (nhit >= 2) stk[(w.sp++) * stride] = c[1];
                    w.cur = c[0];
                } else {
                    w.cur = w.sp ? stk[(--w.sp) * stride] : RT_EMPTY_ROOT;
                }
            } else if (w.cur < 0 && w.leaf == RT_EMPTY_ROOT) {  // park the leaf, go on with the stack
                w.leaf = w.cur;
                w.cur = w.sp ? stk[(--w.sp) * stride] : RT_EMPTY_ROOT;
            }
            const uint64_t parked = __ballot(w.leaf != RT_EMPTY_ROOT);
            const uint64_t nodes = __ballot(w.cur >= 0 && w.cur != RT_EMPTY_ROOT);
            if (parked != 0 && (nodes == 0 || 4 * __popcll(parked) >= PARK * __popcll(__ballot(true)))) {
                if (w.leaf != RT_EMPTY_ROOT) {
                    const bool changed = test_leaf(tris, w.leaf, r, true, skip_slot, h);
                    w.leaf = RT_EMPTY_ROOT;
                    if (changed && (ANY || any_rt)) return kWalkDone;
                }
            }
            if (w.cur == RT_EMPTY_ROOT && w.leaf == RT_EMPTY_ROOT) return kWalkDone;
        } else {
            if (w.cur >= 0) {
                int32_t c[4];
                const int nhit = lane4_children<NF>(nodes4, w.cur, r, w, h.t, c);
                if (nhit > 0) {
                    if (w.sp + nhit - 1 > limit) return kWalkOverflow;
                    if (nhit >= 4) stk[(w.sp++) * stride] = c[3];
                    if (nhit >= 3) stk[(w.sp++) * stride] = c[2];
                    if (nhit >= 2) stk[(w.sp++) * stride] = c[1];
                    w.cur = c[0];
                    continue;
                }
            } else {
                if (test_leaf(tris, w.cur, r, true, skip_slot, h) && (ANY || any_rt)) return kWalkDone;
            }
            if (w.sp == 0) return kWalkDone;
            w.cur = stk[(--w.sp) * stride];
        }
    }
    return kWalkOn;
}

// Per-lane traversal of the 4-wide nodes: every lane walks its own ray (no lane waits on another's
// subtrees, so the longest lane bounds the wave), nearer children first by the lane's own entry
// distances, pending children on a per-lane stack in LDS. `stk` points at this thread's first
// entry; entry k lives at stk[k * stride] (stride = threads per block, so the lanes of a wave touch
// consecutive words at equal depth). Returns false if the ray needed more than `limit` (at most
// kLaneStack) pending entries (3 per level of the 4-wide tree: up to 93): the hit is then invalid
// and the caller traces the ray again with a traversal whose stack always suffices.
template <bool ANY>
RT_DEV bool traverse_lane4(const float4* __restrict__ nodes4, const float4* __restrict__ tris,
                           int32_t root, const Ray& r, int32_t skip_slot, Hit& h, int32_t* stk,
                           int stride, int limit, bool fresh = true) {
    if (fresh) {  // else h holds a hit of this ray found earlier: it bounds the walk
        h.t = r.tmax;
        h.u = 0.0f;
        h.v = 0.0f;
        h.prim = -1;
        h.slot = -1;
    }
    if (root == RT_EMPTY_ROOT) return true;
    LaneWalk w;
    walk_start(r, root, w);
    return lane4_steps<ANY, RT_PARK_RENDER>(nodes4, tris, r, skip_slot, h, w, stk, stride, limit, 0x7FFFFFFF) !=
           kWalkOverflow;
}

// Primary ray direction of pixel (col,row) (rendering/_core.py:528-548 look_at/perspective,
// rendering/_raster.py:126-129 viewport with the y flip, :298-299 pixel centre).
RT_DEV void camera_dir(const rt_camera& cam, uint32_t W, uint32_t H, uint32_t col, uint32_t row,
                       float& dx, float& dy, float& dz) {
    const float nx = ((float)col + 0.5f) / (float)W * 2.0f - 1.0f;
    const float ny = 1.0f - ((float)row + 0.5f) / (float)H * 2.0f;
    const float x = fmaf(nx, cam.right[0], fmaf(ny, cam.up[0], cam.forward[0]));
    const float y = fmaf(nx, cam.right[1], fmaf(ny, cam.up[1], cam.forward[1]));
    const float z = fmaf(nx, cam.right[2], fmaf(ny, cam.up[2], cam.forward[2]));
    const float il = 1.0f / sqrtf(fmaf(x, x, fmaf(y, y, z * z)));
    dx = x * il;
    dy = y * il;
    dz = z * il;
}

// Wave64 compaction: returns this lane's slot in a contiguous range appended to *counter by
// the active lanes with `pred` set (one atomic per wave).
RT_DEV uint32_t wave_append(uint32_t* counter, bool pred) {
    const uint64_t mask = __ballot(pred);
    const uint32_t lo = (uint32_t)mask, hi = (uint32_t)(mask >> 32);
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (mask != 0 && (int)__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader < 0 ? 0 : leader);
    return base + rank;
}

}  // namespace rt

// ---- host-side error plumbing (rt_api.hip) ----
namespace rt {
int set_error(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);
int check_launch(const char* what);
int device_cu_count();
rt_tuning current_tuning();
// the call's tuning: *in validated, or the process default when in is NULL
int resolve_tuning(const rt_tuning* in, rt_tuning& out, const char* what);
}  // namespace rt
