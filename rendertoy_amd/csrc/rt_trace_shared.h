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
