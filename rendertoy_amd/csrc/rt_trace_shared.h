// rt_trace_shared.h — pieces shared by the trace/render kernels (rt_trace.hip) and the path
// kernels (rt_paths.hip): scene arguments, the three ray traversals behind one call, the
// camera-ray coherence rule, and host-side scene checks.
#pragma once

#include "rt_common.h"

namespace rt {

struct SceneArgs {
    const float4* nodes;
    const float4* nodes4;
    const float4* tris;
    const float4* shade;
    const int32_t* mesh_offsets;
    const int32_t* prim_slot;
    int32_t root;
    uint32_t n_meshes;
};

// Traversal of one ray per lane by one of four algorithms (MODE = RT_TRAVERSAL_*); `pk` = the
// wave's kPacketLds-float4 LDS area of the packet traversal:
//  PACKET  the wave's rays walk one node sequence together (traverse_packet);
//  LANE    each lane walks its own ray over the 4-wide nodes with a per-lane stack of `lane_stack`
//          entries in LDS at `stk` (entries `stk_stride` words apart); a ray that overflows it is
//          traced again by its wave with the packet traversal (128-entry stack, enough for any tree);
//  BINARY  as LANE over the binary nodes (traverse_bin), same stack and fallback;
//  HYBRID  packet traversal for at most `budget` steps; if the wave is not done by then, each lane
//          continues alone from the root with its best hit so far as the bound (as LANE).
// All three return the same hit: the closest hit does not depend on the visiting order, and a
// restart bounded by a hit of the same ray visits every box that could hold a closer one.
// Called by all 64 lanes of a wave with uniform control flow.
template <bool ANY, int MODE>
RT_DEV void trace_block(const SceneArgs& sc, const Ray& r, bool want, int32_t skip_slot, Hit& h,
                        float4* pk, int32_t* stk, int stk_stride, int lane_stack, int budget) {
    static_assert((MODE >= RT_TRAVERSAL_PACKET && MODE <= RT_TRAVERSAL_HYBRID) ||
                      MODE == RT_TRAVERSAL_BINARY, "resolved mode");
    if (MODE == RT_TRAVERSAL_PACKET) {
        traverse_packet<ANY>(sc.nodes4, sc.tris, sc.root, r, want, skip_slot, h, pk);
        return;
    }
    bool walk = want, fresh = true;
    if (MODE == RT_TRAVERSAL_HYBRID) {
        if (traverse_packet<ANY>(sc.nodes4, sc.tris, sc.root, r, want, skip_slot, h, pk, budget)) return;
        walk = want && !(ANY && h.prim >= 0);  // an occluded shadow ray is done
        fresh = false;
    } else {
        h.t = r.tmax;
        h.u = 0.0f;
        h.v = 0.0f;
        h.prim = -1;
        h.slot = -1;
    }
    bool ok = true;
    if (MODE == RT_TRAVERSAL_BINARY) {
        if (walk) ok = traverse_bin<ANY>(sc.nodes, sc.tris, sc.root, r, skip_slot, h, stk, stk_stride, lane_stack);
    } else if (walk) {
        ok = traverse_lane4<ANY>(sc.nodes4, sc.tris, sc.root, r, skip_slot, h, stk, stk_stride,
                                 lane_stack, fresh);
    }
    const bool redo = !ok;
    if (__ballot(redo)) {
        Hit h2;
        traverse_packet<ANY>(sc.nodes4, sc.tris, sc.root, r, redo, skip_slot, h2, pk);
        if (redo) h = h2;
    }
}

constexpr bool uses_lds(int mode) { return mode != RT_TRAVERSAL_PACKET; }

// ---- per-lane walks with refill ----
// Walk lengths of incoherent rays are heavy-tailed (on random rays through a 1.3M-triangle scene
// the median ray visits ~18 nodes, the 99th percentile ~480), so a wave holding one ray per lane
// runs as long as its longest lane while most lanes idle. refill_walks keeps the lanes fed instead:
// - rays 0..n-1 are taken in groups of 64 consecutive ones; of G waves, wave `wid` owns the groups
//   wid, wid + G, wid + 2G, ... (consecutive rays stay together, so coherent buffers keep their
//   coherence, and each wave samples the whole buffer, which evens out the waves' loads);
// - the wave walks in rounds of at most kRefillSteps iterations (one node or one leaf) per lane,
//   and whenever kRefillIdle or more lanes are idle, the idle lanes take the wave's next rays in
//   lane order (one ballot);
// - a lane's walk (bin_steps or lane4_steps, MODE; PARK: lane4_steps' leaf rounds) is the one-ray
//   walk resumed across rounds, so every ray gets the hit it gets alone;
// - a ray whose LDS stack would overflow is listed in LDS (`list`, kRefillList entries); when the
//   list could fill, the wave stops taking rays, finishes those it holds and traces the listed ones
//   again with the packet traversal, 64 at a time, then goes on.
// `load(i, r)` loads ray i into r and returns the triangle slot it must skip (-1: none);
// `finish(fin, i, r, h)` is called by all 64 lanes after every round, with fin set on the lanes
// whose ray i has its final hit h (it may use wave-wide operations). Called by all 64 lanes.
#ifndef RT_REFILL_STEPS
#define RT_REFILL_STEPS 16
#endif
#ifndef RT_REFILL_IDLE
#define RT_REFILL_IDLE 8
#endif
constexpr int kRefillSteps = RT_REFILL_STEPS;       // walk iterations per round
constexpr uint32_t kRefillIdle = RT_REFILL_IDLE;    // idle lanes that trigger a refill
constexpr uint32_t kRefillList = 128;               // overflow list entries per wave

RT_DEV uint32_t lane_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// The wave's place in refill_walks' order: blocks are handed to the 8 XCDs round robin (block b runs
// on XCD b % 8), so waves w = b * waves_per_block + i are renumbered XCD-major; the G waves work on
// G consecutive 64-ray groups at a time, and each XCD then takes a contiguous run of them, which
// keeps a coherent buffer's neighbouring rays in one L2. (Identity unless the grid is a multiple of 8.)
RT_DEV uint32_t refill_slot(uint32_t block, uint32_t wave_in_block, uint32_t waves_per_block, uint32_t n_blocks) {
    if (n_blocks & 7u) return block * waves_per_block + wave_in_block;
    return ((block & 7u) * (n_blocks >> 3) + (block >> 3)) * waves_per_block + wave_in_block;
}

// ---- one ray walked by a whole wave (the last rays of a refill_walks sequence) ----
// The 64 lanes walk the ray r (the same in every lane) together: each step takes up to 16 entries
// from the top of a stack shared by the wave (entry e at at(e), at most `cap`), lane j tests child
// j & 3 of entry j >> 2 against the wave's best t, the entered internal children are pushed
// (ballot compaction) and each entered leaf is tested by its lane; the wave's closest candidate
// (t, then the smaller primitive id) is reduced after each step. The entries are taken in stack
// order, not nearest first: the closest hit does not depend on the visiting order, so the result
// is the one-lane walk's. The stack starts with the root, or with `seeded` entries the caller
// wrote (a one-lane walk's pending nodes and leaves; a leaf entry is tested by one lane). h: the
// ray's best candidate so far on entry (the same in every lane), its closest hit on return (ANY:
// a hit). Returns false, h still a valid bound, if a push would pass `cap` entries; the caller
// then finishes the ray another way.
// busy lanes at the end of a sequence that the wave walks this way (measured on 2.07 M camera-order
// and random rays through the 1.31M-triangle scene: 8 for both; 16 takes 168-202 VGPRs and a wave
// or two per SIMD)
constexpr uint32_t kCoopLanesClosest = 8, kCoopLanesAny = 8;

template <bool ANY, typename At>
RT_DEV bool wave_walk(const float4* __restrict__ nodes4, const float4* __restrict__ tris, int32_t root,
                      const Ray& r, int32_t skip_slot, Hit& h, At at, uint32_t cap, uint32_t seeded = 0) {
    const uint32_t lane = __lane_id(), c = lane & 3u, entry = lane >> 2;
    if (root == RT_EMPTY_ROOT) return true;
    // merges the lanes' candidates (each no worse than h, some equal to it) into h
    auto reduce = [&](const Hit& cand) {
        float t = cand.t;
        int32_t p = cand.prim;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const float ot = __shfl_xor(t, off);
            const int32_t op = __shfl_xor(p, off);
            const bool take = op >= 0 && (p < 0 || ot < t || (ot == t && op < p));
            t = take ? ot : t;
            p = take ? op : p;
        }
        if (p < 0 || !(t < h.t || (t == h.t && p < h.prim))) return;
        const uint64_t win = __ballot(cand.prim == p && cand.t == t);
        const int wl = __ffsll((unsigned long long)win) - 1;
        h.t = t;
        h.prim = p;
        h.u = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cand.u), wl));
        h.v = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cand.v), wl));
        h.slot = __builtin_amdgcn_readlane(cand.slot, wl);
    };
    // this lane's leaf, one triangle at a time (few registers: this runs inside the refill loop)
    auto leaf_one = [&](int32_t ref, Hit& cand) {
        const uint32_t first = leaf_first(ref), count = leaf_count(ref);
        for (uint32_t k = 0; k < count; ++k) {
            const int32_t slot = (int32_t)(first + k);
            const float4* tr = tris + (size_t)slot * 3;
            float tt, uu, vv;
            const float4 ta = tr[0];
            if (slot != skip_slot && intersect_tri(ta, tr[1], tr[2], r, tt, uu, vv) && tt > r.tmin) {
                const int32_t prim = __float_as_int(ta.w);
                if (tt < cand.t || (tt == cand.t && prim < cand.prim)) {
                    cand.t = tt;
                    cand.u = uu;
                    cand.v = vv;
                    cand.prim = prim;
                    cand.slot = slot;
                }
            }
        }
    };
    if (root < 0) {  // the whole tree is one leaf
        Hit cand = h;
        if (lane == 0) leaf_one(root, cand);
        reduce(cand);
        return true;
    }
    const float idx = safe_inv(r.dx), idy = safe_inv(r.dy), idz = safe_inv(r.dz);
    const float oxx = r.ox * idx, oyy = r.oy * idy, ozz = r.oz * idz;
    uint32_t sp = seeded;  // entries 0..seeded-1: the ray's pending nodes and leaves, pushed by the caller
    if (!seeded) {
        if (lane == 0) *at(0u) = root;
        sp = 1;
    }
    while (sp > 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t k = sp < 16u ? sp : 16u, base = sp - k;
        const int32_t node = entry < k ? *at(base + entry) : RT_EMPTY_ROOT;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        sp = base;
        bool push = false, leaf = false;
        int32_t cref = RT_EMPTY_ROOT;
        if (node < 0 && node != RT_EMPTY_ROOT) {  // a pending leaf (from a seeded stack): lane c = 0 tests it
            leaf = c == 0u;
            cref = node;
        } else if (node != RT_EMPTY_ROOT) {
            const float* nf = (const float*)(nodes4 + (size_t)node * 8);
            const float clx = nf[c], chx = nf[4 + c], cly = nf[8 + c], chy = nf[12 + c], clz = nf[16 + c],
                        chz = nf[20 + c];
            cref = __float_as_int(nf[24 + c]);
            const float x0 = fmaf(clx, idx, -oxx), x1 = fmaf(chx, idx, -oxx);
            const float y0 = fmaf(cly, idy, -oyy), y1 = fmaf(chy, idy, -oyy);
            const float z0 = fmaf(clz, idz, -ozz), z1 = fmaf(chz, idz, -ozz);
            const float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), r.tmin));
            const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1)) * kTfarScale;
            const bool hk = (cref != RT_EMPTY_ROOT) & (tn <= fminf(tf, h.t));
            push = hk & (cref >= 0);
            leaf = hk & (cref < 0);
        }
        const uint64_t pm = __ballot(push);
        const uint32_t np = (uint32_t)__popcll(pm);
        if (sp + np > cap) return false;
        if (push) *at(sp + lane_rank(pm)) = cref;
        sp += np;
        if (__ballot(leaf)) {
            Hit cand = h;
            if (leaf) leaf_one(cref, cand);
            reduce(cand);
            if (ANY && h.prim >= 0) return true;
        }
    }
    return true;
}

template <bool ANY, int MODE, int PARK, bool NF = false, bool COOP = true, typename Load, typename Finish>
RT_DEV void refill_walks(const SceneArgs& sc, uint64_t n, uint32_t n_waves, uint32_t wid, int32_t* stk,
                         int stk_stride, int lane_stack, uint32_t* list, float4* pk, Load&& load,
                         Finish&& finish) {
    static_assert(MODE == RT_TRAVERSAL_LANE || MODE == RT_TRAVERSAL_BINARY, "per-lane walk");
    // position p of the wave's sequence -> ray index (group p / 64 of the wave, offset p % 64)
    auto ray_at = [&](uint64_t p) -> uint64_t { return (((p >> 6) * n_waves + wid) << 6) | (p & 63u); };
    uint64_t cursor = 0;             // wave-uniform
    bool more = ray_at(0) < n;       // wave-uniform: rays left in the wave's sequence
    const uint32_t lane = __lane_id();
    while (more) {
        uint32_t listed = 0;  // wave-uniform
        bool busy = false;
        uint32_t ri = 0;      // the lane's ray while busy
        int32_t skip = -1;
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
        for (;;) {
            const uint64_t idle = __ballot(!busy);
            const uint32_t n_idle = (uint32_t)__popcll(idle);
            if (more && listed + 64u <= kRefillList && (n_idle >= kRefillIdle || n_idle == 64u)) {
                if (!busy) {
                    const uint64_t i = ray_at(cursor + lane_rank(idle));
                    if (i < n) {
                        ri = (uint32_t)i;
                        skip = load(ri, r);
                        h.t = r.tmax;
                        h.u = h.v = 0.0f;
                        h.prim = h.slot = -1;
                        walk_start(r, sc.root, w);
                        busy = true;
                    }
                }
                cursor += n_idle;
                more = ray_at(cursor) < n;  // later positions map to larger indices
            }
            if (__ballot(busy) == 0) break;
            int st = kWalkOn;
            if (busy) {
                if (sc.root == RT_EMPTY_ROOT)
                    st = kWalkDone;
                else if (MODE == RT_TRAVERSAL_BINARY)
                    st = bin_steps<ANY>(sc.nodes, sc.tris, r, skip, h, w, stk, stk_stride, lane_stack, kRefillSteps);
                else
                    st = lane4_steps<ANY, PARK, NF>(sc.nodes4, sc.tris, r, skip, h, w, stk, stk_stride, lane_stack,
                                          kRefillSteps);
            }
            finish(st == kWalkDone, ri, r, h);
            if (st != kWalkOn) busy = false;
            const uint64_t over = __ballot(st == kWalkOverflow);
            if (over) {  // refills stop while listed > kRefillList - 64: listed stays <= kRefillList
                if (st == kWalkOverflow) list[listed + lane_rank(over)] = ri;
                listed += (uint32_t)__popcll(over);
            }
            // The sequence's last rays: once nothing is left to take and at most kCoop lanes walk,
            // the whole wave walks each of those rays in turn (wave_walk) on the wave's lane stacks,
            // from where its one-lane walk stopped (its stack entries, current node and parked leaf,
            // read from all the lanes before the first wave walk takes the stacks). A ray whose
            // shared stack would overflow goes to the packet retrace like an overflow. A sequence's
            // slowest rays set the launch's floor: 32 K random rays took 1.13 ms (any hit 0.92) one
            // lane each, 0.61 (0.38) this way (0.79 / 0.38 restarted from the root, 0.67 closest
            // with the wave walking a sequence's last 4 rays instead of 8).
            constexpr uint32_t kCoop = ANY ? kCoopLanesAny : kCoopLanesClosest;
            if (COOP && MODE == RT_TRAVERSAL_LANE && !more) {
                uint64_t bm = __ballot(busy);
                if (bm != 0 && (uint32_t)__popcll(bm) <= kCoop && listed + kCoop <= kRefillList) {
                    int32_t* stk0 = stk - lane;  // lane 0's column of the wave's stacks
                    auto at = [&](uint32_t e) -> int32_t* { return stk0 + (size_t)(e >> 6) * stk_stride + (e & 63u); };
                    // the rays' pending work (stack entries, current node, parked leaf), read before
                    // the first wave walk takes the stacks: entry j of ray k in lane j's pe[k]
                    int32_t pe[kCoop], pcur[kCoop], pleaf[kCoop];
                    uint32_t psp[kCoop];
                    int pb[kCoop];
#pragma unroll
                    for (uint32_t k = 0; k < kCoop; ++k) {
                        pb[k] = bm ? __ffsll((unsigned long long)bm) - 1 : -1;
                        bm &= bm ? bm - 1 : 0ull;
                        const int bk = pb[k] < 0 ? 0 : pb[k];
                        psp[k] = pb[k] < 0 ? 0u : (uint32_t)__builtin_amdgcn_readlane(w.sp, bk);
                        pcur[k] = pb[k] < 0 ? RT_EMPTY_ROOT : __builtin_amdgcn_readlane(w.cur, bk);
                        pleaf[k] = pb[k] < 0 ? RT_EMPTY_ROOT : __builtin_amdgcn_readlane(w.leaf, bk);
                        pe[k] = lane < psp[k] ? stk0[(size_t)lane * stk_stride + (uint32_t)bk] : RT_EMPTY_ROOT;
                    }
#pragma unroll
                    for (uint32_t k = 0; k < kCoop; ++k) {
                        const int b = pb[k];
                        if (b < 0) break;
                        Ray rb;
                        rb.ox = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r.ox), b));
                        rb.oy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r.oy), b));
                        rb.oz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r.oz), b));
                        rb.dx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r.dx), b));
                        rb.dy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r.dy), b));
                        rb.dz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r.dz), b));
                        rb.tmin = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r.tmin), b));
                        rb.tmax = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r.tmax), b));
                        Hit hb;
                        hb.t = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(h.t), b));
                        hb.u = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(h.u), b));
                        hb.v = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(h.v), b));
                        hb.prim = __builtin_amdgcn_readlane(h.prim, b);
                        hb.slot = __builtin_amdgcn_readlane(h.slot, b);
                        const int32_t skb = __builtin_amdgcn_readlane(skip, b);
                        // the shared stack starts with the ray's pending work, so its walk goes on
                        // where the lane's stopped (a ray with none left is done)
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                        __builtin_amdgcn_wave_barrier();
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                        if (lane < psp[k]) *at(lane) = pe[k];
                        uint32_t seeded = psp[k];
                        if (pcur[k] != RT_EMPTY_ROOT) {
                            if (lane == 0) *at(seeded) = pcur[k];
                            ++seeded;
                        }
                        if (pleaf[k] != RT_EMPTY_ROOT) {
                            if (lane == 0) *at(seeded) = pleaf[k];
                            ++seeded;
                        }
                        const bool ok = seeded == 0u || wave_walk<ANY>(sc.nodes4, sc.tris, sc.root, rb, skb, hb, at,
                                                                       64u * (uint32_t)lane_stack, seeded);
                        if (ok) {
                            finish((int)lane == b, ri, r, hb);
                        } else {
                            if ((int)lane == b) list[listed] = ri;
                            listed += 1u;
                        }
                        if ((int)lane == b) busy = false;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    break;  // every lane is idle and nothing is left to take
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the list's stores before its loads
        for (uint32_t k = 0; k < listed; k += 64u) {
            const bool want = k + lane < listed;
            const uint32_t i = want ? list[k + lane] : 0u;
            Ray rr = r;  // the placeholder's fields for lanes without a listed ray
            int32_t sk = -1;
            if (want) sk = load(i, rr);
            Hit h2;
            traverse_packet<ANY>(sc.nodes4, sc.tris, sc.root, rr, want, sk, h2, pk);
            finish(want, i, rr, h2);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // its loads before the next stores
    }
}

// RT_TRAVERSAL_AUTO for camera rays: the packet traversal when the 8x8 tile of a wave spans at most
// kCoherentTile in the image plane at unit distance (8 pixels x 2 tan(fov/2) / height, or the
// width's counterpart), where its rays share most of their nodes; per-lane beyond.
constexpr float kCoherentTile = 0.009f;

__host__ __device__ inline bool coherent_tiles(const rt_camera& c, uint32_t W, uint32_t H) {
    const float up = sqrtf(c.up[0] * c.up[0] + c.up[1] * c.up[1] + c.up[2] * c.up[2]);
    const float rt = sqrtf(c.right[0] * c.right[0] + c.right[1] * c.right[1] + c.right[2] * c.right[2]);
    const float fy = H ? 16.0f * up / (float)H : 0.0f, fx = W ? 16.0f * rt / (float)W : 0.0f;
    return (fx > fy ? fx : fy) <= kCoherentTile;
}

// Camera-ray traversal with AUTO resolved per wave from `coherent` (uniform).
template <int PM>
RT_DEV void trace_primary(const SceneArgs& sc, const Ray& r, bool want, Hit& h, float4* pk,
                          int32_t* stk, int stk_stride, int lane_stack, int budget, bool coherent) {
    if constexpr (PM == RT_TRAVERSAL_AUTO) {
        if (coherent)
            trace_block<false, RT_TRAVERSAL_PACKET>(sc, r, want, -1, h, pk, stk, stk_stride, lane_stack,
                                                    budget);
        else
            trace_block<false, RT_TRAVERSAL_LANE>(sc, r, want, -1, h, pk, stk, stk_stride, lane_stack,
                                                  budget);
    } else {
        trace_block<false, PM>(sc, r, want, -1, h, pk, stk, stk_stride, lane_stack, budget);
    }
}

inline SceneArgs scene_args(const rt_scene* s) {
    SceneArgs a;
    a.nodes = (const float4*)s->nodes;
    a.nodes4 = (const float4*)s->nodes4;
    a.tris = (const float4*)s->tris;
    a.shade = (const float4*)s->shade;
    a.mesh_offsets = s->mesh_offsets;
    a.prim_slot = s->prim_slot;
    a.root = s->root;
    a.n_meshes = s->n_meshes;
    return a;
}

inline int check_scene(const rt_scene* s, const char* what) {
    if (!s) return set_error(RT_EINVAL, std::string(what) + ": scene is NULL");
    if (s->root != RT_EMPTY_ROOT && (!s->nodes || !s->nodes4 || !s->tris || !s->shade || !s->mesh_offsets))
        return set_error(RT_EINVAL, std::string(what) + ": scene arrays are NULL");
    if (s->root != RT_EMPTY_ROOT && s->n_tris == 0)
        return set_error(RT_EINVAL, std::string(what) + ": scene root set but n_tris is 0");
    return RT_OK;
}

// The camera-ray traversal a kernel with PACKET / LANE / HYBRID variants runs for `mode`: AUTO
// decided by the coherence rule, QUAD (implemented by rt_trace_camera only) as PACKET.
inline int resolve_primary(int mode, const rt_camera& cam, uint32_t W, uint32_t H) {
    if (mode == RT_TRAVERSAL_QUAD) return RT_TRAVERSAL_PACKET;
    if (mode != RT_TRAVERSAL_AUTO) return mode;
    return coherent_tiles(cam, W, H) ? RT_TRAVERSAL_PACKET : RT_TRAVERSAL_LANE;
}

}  // namespace rt
