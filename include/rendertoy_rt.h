/*
 * rendertoy_rt.h — C ABI of librtx_hip.so, the MI355X (gfx950) ray-casting engine behind
 * the rendertoy raycaster API.
 *
 * The reference raycaster is a Python class whose device work was meant to run as pyopencl
 * kernels dispatched through `kernel_main` (rendering/_core.py:247-299). Its two entry points are
 * stubs: `Raycaster._build_ads` (rendering/_raycaster.py:30-33) and `Raycaster.ray_cast`
 * (rendering/_raycaster.py:35-36). This header declares the native entry points that the Python
 * mirror of that class (rendertoy_amd/_raycaster.py) binds with ctypes. A maintainer wiring the
 * reference itself would bind the same symbols from rendering/_raycaster.py.
 *
 * Conventions
 *  - Every pointer named `*_dev` or documented as "device" is a device address (hipMalloc'd
 *    memory, e.g. torch.Tensor.data_ptr()). The library never allocates or frees device memory:
 *    the caller passes every buffer, including workspace.
 *  - Every call is asynchronous on the caller's stream (`stream` is a hipStream_t; NULL means
 *    the default stream) unless documented otherwise.
 *  - Return value: RT_OK (0) or a negative RT_E* code; rt_last_error() returns a thread-local
 *    message for the last failure on the calling thread.
 *  - Triangle identity: global primitive id = offset of the mesh (sum of the triangle counts of
 *    the meshes before it) + triangle index within the mesh; the hit reports (mesh, index) as
 *    rendering/_raycaster.py:16-22 (BVH_Triangle.index / .mesh) defines them.
 */
#ifndef RENDERTOY_RT_H
#define RENDERTOY_RT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RT_ABI_VERSION 10

enum {
    RT_OK = 0,
    RT_EINVAL = -1,    /* bad argument (null pointer, size mismatch, misaligned buffer) */
    RT_EHIP = -2,      /* a HIP runtime call failed (message holds hipGetErrorString) */
    RT_ECAPACITY = -3, /* a size exceeds what the kernels support */
    RT_ENOTSUP = -4    /* unsupported option */
};

/* Byte offset marking an absent per-vertex attribute in rt_mesh_desc. */
#define RT_ABSENT 0xFFFFFFFFu

/* One input mesh. Mirrors rendering/_modeling.py:31-34 (Mesh(vertices, indices)) over the
 * 80-byte MeshVertex record of rendering/_modeling.py:22-28 (P@0, N@16, C@32, T@48, B@64). */
typedef struct rt_mesh_desc {
    const void* vertices;     /* device: n_vertices records of vertex_stride bytes */
    const int32_t* indices;   /* device: 3*n_tris vertex indices, or NULL: vertices 3t..3t+2 */
    uint32_t n_vertices;
    uint32_t n_tris;          /* indices ? (index count)/3 : n_vertices/3 (rendering/_raycaster.py:33) */
    uint32_t vertex_stride;   /* bytes between vertex records (80 for MeshVertex) */
    uint32_t pos_offset;      /* byte offset of float3 P */
    uint32_t nrm_offset;      /* byte offset of float3 N, or RT_ABSENT (geometric normal used) */
    uint32_t uv_offset;       /* byte offset of float2 C (texture coordinate), or RT_ABSENT */
} rt_mesh_desc;

/* A built scene: device pointers owned by the caller, sizes from rt_scene_sizes_for().
 * Node references: >= 0 an internal node id; < 0 a leaf of count = ((~ref) >> 29) + 1 consecutive
 * triangle slots starting at (~ref) & 0x1FFFFFFF (subtrees of <= 4 triangles are leaves);
 * RT_EMPTY_ROOT: no triangles. */
typedef struct rt_scene {
    uint32_t n_tris;          /* at most 2^29 - 1 */
    uint32_t n_meshes;
    int32_t root;
    uint32_t _reserved;
    void* nodes;              /* device, max(n_tris-1,1) * 64 bytes: binary nodes (4 x float4) */
    void* nodes4;             /* device, max(n_tris-1,1) * 128 bytes: 4-wide nodes (8 x float4) */
    void* tris;               /* device, n_tris * 48 bytes (v0|prim, e1, e2 as float4) */
    void* shade;              /* device, n_tris * 64 bytes (N0,N1,N2,UV0,UV1,UV2,mesh) */
    int32_t* mesh_offsets;    /* device, n_meshes+1 prefix sums of triangle counts */
    float* bounds;            /* device, 8 floats: scene AABB lo.xyz, 0, hi.xyz, 0 */
    int32_t* prim_slot;       /* device, n_tris: the triangle-record slot of each global primitive id */
} rt_scene;

#define RT_EMPTY_ROOT ((int32_t)0x7FFFFFFF)

/* Per-call byte sizes of the scene arrays and of the build workspace for n_tris triangles. */
typedef struct rt_scene_sizes {
    uint64_t nodes_bytes, nodes4_bytes, tris_bytes, shade_bytes, bounds_bytes, workspace_bytes;
    uint64_t prim_slot_bytes;
} rt_scene_sizes;

/* Rays as field pointers plus one byte stride shared by all fields: SoA arrays use stride 4,
 * the AoS Ray record {float3 origin @0; float3 direction @16} uses stride 32. tmin/tmax may be
 * NULL, then tmin_default/tmax_default apply to every ray. */
typedef struct rt_rays {
    const float *ox, *oy, *oz, *dx, *dy, *dz;
    const float *tmin, *tmax;
    uint32_t stride;
    float tmin_default, tmax_default;
} rt_rays;

/* Closest-hit output fields (device), same stride rule as rt_rays. Misses store t=+inf,
 * u=v=0, prim=index=mesh=-1. Any pointer may be NULL (field not written). */
typedef struct rt_hits {
    float *t, *u, *v;
    int32_t *prim, *index, *mesh;
    uint32_t stride;
} rt_hits;

/* Pinhole camera of rendering/_core.py:528-548 (look_at + perspective, vertical fov) reduced to
 * the four vectors the ray generator uses: the primary direction of pixel (col,row) of a W x H
 * image is normalize(nx*right + ny*up + forward) with nx = 2(col+.5)/W - 1, ny = 1 - 2(row+.5)/H,
 * right = xaxis * aspect*tan(fov/2), up = yaxis * tan(fov/2), forward = zaxis. */
typedef struct rt_camera {
    float eye[3];
    float right[3];
    float up[3];
    float forward[3];
} rt_camera;

/* RGBA8 texture, row-major rows of `width` texels, row index = v (rendering/_core.py:94-96). */
typedef struct rt_texture {
    const uint8_t* texels;    /* device, width*height*4 bytes */
    uint32_t width, height;
} rt_texture;

enum {
    RT_RENDER_SHADOWS = 1,    /* trace one shadow ray toward the light per lit hit */
    RT_RENDER_BILINEAR = 2    /* bilinear texture filter (else point sampling as sample2D) */
};

/* Shading of rendering/_core.py:94-96 and tutorials/lesson09_texture_mapping.py:73,93-94:
 * rgb = albedo(uv) * (ambient + V * max(0, N.L)); misses store `background`. */
typedef struct rt_render_params {
    uint32_t width, height;
    uint32_t flags;
    float light[3];           /* unit direction toward the light */
    float ambient;            /* 0.2 in the reference shaders */
    float background[3];
    float shadow_tmin;        /* shadow-ray start offset along the light direction */
    const rt_texture* textures_dev;  /* device array of rt_texture, or NULL */
    const int32_t* mesh_texture_dev; /* device, per mesh: texture slot or -1 (albedo 1) */
    uint32_t n_textures;
} rt_render_params;

/* Output of rt_render. rgb is width*height*3 float32 (row-major, row 0 = top). The hit fields
 * are optional (NULL = not written); `stats` (optional, device, 4 uint32) receives
 * {primary hits, shadow rays traced, shadow rays occluded, 0}. */
typedef struct rt_render_out {
    float* rgb;
    float* t;
    int32_t* prim;
    uint32_t* stats;
} rt_render_out;

/* Traversal algorithms of the camera and render kernels. Both return the same hits (the closest
 * hit does not depend on the visiting order); they differ in speed only. */
enum {
    RT_TRAVERSAL_PACKET = 0, /* a wave's 64 rays walk one node sequence (scalar node fetches) */
    RT_TRAVERSAL_LANE = 1,   /* each lane walks its own ray (per-lane stack in LDS) */
    RT_TRAVERSAL_HYBRID = 2, /* packet for at most packet_budget steps, then per lane from the root
                                bounded by the best hit so far (for rays that diverge) */
    RT_TRAVERSAL_AUTO = 3,   /* camera rays only: PACKET when an 8x8 pixel tile spans at most 0.009
                                of the image plane at unit distance (16 tan(fov/2) / height, e.g.
                                45 deg at >= 768 rows), else LANE; decided per image */
    RT_TRAVERSAL_QUAD = 4,   /* camera rays: packets of 16 rays (4x4 pixels), 4 lanes per ray, each
                                lane testing one child box / one leaf triangle (the default) */
    RT_TRAVERSAL_BINARY = 5  /* ray buffers (rt_trace_closest / rt_trace_any): each lane walks its
                                own ray over the binary nodes, nearer child first, per-lane stack in
                                LDS (the default for ray buffers) */
};
#define RT_LANE_STACK_MAX 32

/* Launch choices of the build, trace and render entry points (no counterpart in the reference,
 * whose kernels take a fixed 32-thread group, _core.py:286). Every entry point that uses them takes
 * a `const rt_tuning*`: the caller's choices for that call, or NULL for the process default set by
 * rt_set_tuning (so two callers on two streams never share mutable state). */
typedef struct rt_tuning {
    int32_t primary;          /* RT_TRAVERSAL_* for camera rays (default QUAD) */
    int32_t shadow;           /* RT_TRAVERSAL_* for rt_render's shadow rays (default LANE; not used
                                 with a QUAD primary); not QUAD, which is for camera rays */
    uint32_t super_tile;      /* blocks are mapped to XCDs in super-tiles of this many 16x16-pixel
                                 blocks per side; 0 = row-major block order (default) */
    uint32_t lane_stack;      /* per-lane stack entries, 1..RT_LANE_STACK_MAX (default 32); a ray
                                 needing more is traced again by its wave with the packet traversal */
    uint32_t packet_budget;   /* RT_TRAVERSAL_HYBRID: packet steps before the per-lane walk (256) */
    uint32_t builder;         /* RT_BUILD_* hierarchy of rt_scene_build (default LBVH) */
    int32_t query;            /* RT_TRAVERSAL_LANE (per-lane walk of the 4-wide nodes, default) or
                                 RT_TRAVERSAL_BINARY (of the binary nodes) for rt_trace_closest /
                                 rt_trace_any; both keep their stacks in LDS (lane_stack entries per
                                 lane) and trace a ray that needs more again with the packet
                                 traversal */
    uint32_t handoff_steps;   /* must be 0: reserved for a hand-off of the slowest camera packets' walks
                                 to other waves, which is switched off (its first runs did not finish) */
    uint32_t item_steps;      /* reserved with handoff_steps: 1..2^20, default 32 */
} rt_tuning;

/* Hierarchy builders of rt_scene_build; both start from the same Morton order and give the same
 * hits. The PLOC tree (locally ordered clustering by surface area) needs ~30% fewer traversal
 * steps per packet and ~17% fewer instructions than the LBVH radix tree, and takes ~3.7x longer
 * to build; frames whose time is set by their slowest tiles gain little from it. A PLOC tree with
 * a triangle more than 62 internal levels deep (the depth the traversal stacks are sized for, and
 * the bound of the radix tree) is replaced by the LBVH; checking this makes a PLOC build wait for
 * its stream once. */
enum {
    RT_BUILD_LBVH = 0, /* Karras radix tree over the Morton order, leaves of <= 4 consecutive slots */
    RT_BUILD_PLOC = 1  /* PLOC (Meister & Bittner 2018), search radius 16, leaves of <= 4 triangles */
};

/* ---- library ---- */
int rt_abi_version(void);
/* Identity of the sources the library was built from (a hash of its source files), so a loader can
 * refuse a library older than the sources next to it. */
const char* rt_build_id(void);
const char* rt_last_error(void);
int rt_get_tuning(rt_tuning* out);            /* the process default */
int rt_set_tuning(const rt_tuning* tuning);   /* sets the process default; RT_EINVAL if a field is
                                                 out of range (as for a per-call rt_tuning) */

/* ---- scene build: rendering/_raycaster.py:26-33 (Raycaster.__init__ / _build_ads) ---- */

/* Fills `out` with the scene-array and workspace byte sizes for n_tris triangles. */
int rt_scene_sizes_for(uint32_t n_tris, rt_scene_sizes* out);

/* Builds the BVH of all meshes' triangles on the device: per-triangle AABB and centroid,
 * 30-bit Morton codes over the centroid bounds, stable LSD radix sort of (code, primitive), then
 * the hierarchy of rt_tuning.builder: the Karras radix tree with a bottom-up AABB refit (LBVH), or
 * PLOC clustering over the Morton order. `scene` must hold n_tris/n_meshes and the device arrays
 * sized by rt_scene_sizes_for(); on return scene->root is set. `keys_out`/`order_out` (device,
 * n_tris uint32 each, may be NULL) receive the sorted Morton codes and the primitive of each leaf
 * slot (for LBVH the Morton order itself). */
int rt_scene_build(const rt_mesh_desc* meshes, uint32_t n_meshes, rt_scene* scene,
                   void* workspace_dev, uint64_t workspace_bytes,
                   uint32_t* keys_out, uint32_t* order_out, const rt_tuning* tuning, void* stream);

/* Export of a built scene in the reference's record types (rendering/_raycaster.py:8-22):
 *  BVH_AABB     {float3 min @0; float3 max @16; int count @32; int elements @36} (48 B),
 *  BVH_Triangle {float3 v0 @0; float3 v1 @16; float3 v2 @32; int index @48; int mesh @52} (64 B).
 * `aabb_out` receives 2*n_tris - 1 nodes (1 for an empty scene): node 0 is the root with the scene
 * bounds; an internal node has count 0 and its two children at `elements` and `elements` + 1; a
 * leaf has count 1..4 and its triangles at `elements` .. `elements` + count - 1 of `tris_out`
 * (n_tris records in leaf order; index = triangle within its mesh, mesh = position in the mesh
 * list, as the hit identity). Boxes are the hierarchy's conservative (padded) child boxes, so every
 * triangle lies inside each box on its path; v0, v1, v2 are read from the meshes the scene was
 * built from (the same descriptors). An empty scene exports one node with count 0, elements -1, min = +inf,
 * max = -inf. Both outputs must be 16-byte aligned. `n_nodes_out` (device, one uint32, may be
 * NULL) receives the number of node records written, 2 L - 1 for L leaves. */
int rt_scene_export(const rt_scene* scene, const rt_mesh_desc* meshes, uint32_t n_meshes,
                    void* aabb_out, void* tris_out, uint32_t* n_nodes_out, void* stream);

/* ---- queries: rendering/_raycaster.py:35-36 (Raycaster.ray_cast) ---- */

/* Closest hit of n rays (two-sided Moller-Trumbore, hit interval (tmin, tmax), ties in t go to
 * the smaller global primitive id). */
int rt_trace_closest(const rt_scene* scene, const rt_rays* rays, uint32_t n,
                     const rt_hits* hits, const rt_tuning* tuning, void* stream);

/* Any hit (occlusion) of n rays in (tmin, tmax); occluded[i] = 1 if some triangle is hit. */
int rt_trace_any(const rt_scene* scene, const rt_rays* rays, uint32_t n,
                 uint8_t* occluded_dev, const rt_tuning* tuning, void* stream);

/* Camera rays of a W x H image written as rays (SoA, stride 4), for inspection and tests. */
int rt_generate_rays(const rt_camera* cam, uint32_t width, uint32_t height,
                     float* ox, float* oy, float* oz, float* dx, float* dy, float* dz,
                     void* stream);

/* Closest hit of the primary rays of a W x H image (ray generation fused into the traversal);
 * hit record of pixel (col,row) at index row*W + col. `workspace_dev` needs
 * rt_trace_camera_workspace_bytes(W, H) bytes (currently 0: NULL is accepted). */
uint64_t rt_trace_camera_workspace_bytes(uint32_t width, uint32_t height);
int rt_trace_camera(const rt_scene* scene, const rt_camera* cam, uint32_t width, uint32_t height,
                    const rt_hits* hits, void* workspace_dev, uint64_t workspace_bytes,
                    const rt_tuning* tuning, void* stream);

/* Full render in one kernel: per 8x8 pixel tile, ray generation, closest-hit trace, FP64 hit
 * refinement, Lambert with interpolated normals, one shadow ray toward the light per lit hit,
 * RGBA8 texture fetch, RGB store. `workspace_dev` takes rt_render_workspace_bytes(width*height)
 * bytes (32 KiB): with the QUAD traversal a render that returns counts keeps per-slot partial
 * counts there; NULL (or fewer bytes) is accepted, the counts are then summed per 4-wave workgroup
 * (slower). Two calls in flight must not share a workspace. With the QUAD camera-ray traversal (the default) the
 * primary and the shadow rays both take the quad-lane traversal (16-ray packets over 4x4 pixels, 4
 * lanes per ray) and `tuning->shadow` is not used; with the others the primary rays take the
 * chosen traversal and the shadow rays `tuning->shadow`. */
uint64_t rt_render_workspace_bytes(uint32_t n_pixels);
int rt_render(const rt_scene* scene, const rt_camera* cam, const rt_render_params* params,
              const rt_render_out* out, void* workspace_dev, uint64_t workspace_bytes,
              const rt_tuning* tuning, void* stream);

/* ---- batched render: many independent (scene, camera, image) jobs in one persistent launch ----
 * No reference counterpart (the reference renders one image per call). One job = one rt_render
 * call's arguments; the array lives in device memory so that a captured graph can be replayed
 * after the caller rewrites it. */
typedef struct rt_render_job {
    rt_scene scene;
    rt_camera cam;
    rt_render_params params;
    rt_render_out out;
} rt_render_job;

typedef struct rt_batch_config {
    uint32_t waves_per_cu;    /* persistent waves per compute unit, 1..32 */
    uint32_t tiles_per_grab;  /* consecutive pixel tiles (8x8; 4x4 with a QUAD primary) a wave
                                 takes per queue pop, 1..64 */
    int32_t primary;          /* RT_TRAVERSAL_* for the camera rays (AUTO: decided per job; QUAD:
                                 4x4-pixel tiles, shadow rays by the quad-lane traversal too) */
    int32_t shadow;           /* RT_TRAVERSAL_* for the shadow rays (ignored with a QUAD primary) */
    uint32_t lane_stack;      /* per-lane stack entries, 1..RT_LANE_STACK_MAX */
    uint32_t packet_budget;   /* RT_TRAVERSAL_HYBRID packet steps, >= 1 */
} rt_batch_config;

#define RT_BATCH_MAX_JOBS 4096u
#define RT_BATCH_MAX_JOB_PIXELS (1u << 24)

/* Workspace of rt_render_batch for n_jobs jobs (work counter, per-job tile prefix, status). */
uint64_t rt_render_batch_workspace_bytes(uint32_t n_jobs);

/* Renders every job exactly as rt_render would (bit-identical output), in two kernels on `stream`:
 * a prepare kernel (per-job 8x8-tile counts and their prefix sum, work counter and every job's
 * stats zeroed) and one persistent kernel whose waves pull tiles of any job from a global queue.
 * Graph-capturable (no host synchronisation, no memset nodes). `jobs_dev`: device array of n_jobs
 * records (1..RT_BATCH_MAX_JOBS). A job that cannot be rendered is skipped and sets a bit of the
 * status word (rt_render_batch_status): 1 = more than RT_BATCH_MAX_JOB_PIXELS pixels, 2 = a NULL
 * rgb output, texture table (with n_textures > 0) or scene array (non-empty scene). Jobs may share
 * scenes and textures; their output buffers must not overlap. */
int rt_render_batch(const rt_render_job* jobs_dev, uint32_t n_jobs, const rt_batch_config* cfg,
                    void* workspace_dev, uint64_t workspace_bytes, void* stream);

/* Device address of the uint32 status word inside a batch workspace (0 = every job rendered). */
const uint32_t* rt_render_batch_status(const void* workspace_dev);

/* rt_render_batch without a workspace argument and without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The batch workspace is a file-scope device array
 * of the library sized for RT_BATCH_MAX_JOBS jobs (16,644 bytes, one copy per device, held until
 * the process ends); more jobs return RT_ECAPACITY. Output is bit-identical to rt_render_batch.
 * Every call on one device shares that array: calls whose launches overlap on one device (e.g. on
 * two streams) race; issue them on one stream or wait between them. The status word is not
 * readable (a skipped job leaves its outputs unwritten). */
int rt_render_batch_static(const rt_render_job* jobs_dev, uint32_t n_jobs, const rt_batch_config* cfg,
                           void* stream);

/* ---- incoherent rays: diffuse paths (no reference counterpart) ----
 * spp jittered camera rays per pixel, each followed by up to `bounces` cosine-weighted diffuse
 * bounces off the geometric normal (facing the incoming ray), lit by a uniform sky: a sample's
 * radiance is sky * albedo^k if its path escapes after k bounces, 0 if it is still on a surface
 * after the last one; the pixel is the mean of its samples. Random numbers: Philox4x32-10 with
 * key (seed, frame) and counter (pixel, sample, depth, 0); depth 0 jitters the camera ray, depth k
 * draws bounce k. Rays that hit are compacted into a queue per depth (wave ballot, one atomic per
 * wave). Deterministic: the output does not depend on queue order. */
#define RT_PATHS_MAX_BOUNCES 16u

typedef struct rt_paths_params {
    uint32_t width, height;
    uint32_t spp;             /* samples per pixel, 1..1024 */
    uint32_t bounces;         /* 0..RT_PATHS_MAX_BOUNCES */
    uint32_t seed;            /* Philox key word 0 */
    uint32_t frame;           /* Philox key word 1 */
    float albedo;             /* diffuse reflectance of every surface */
    float sky;                /* radiance of the sky */
    float ray_tmin;           /* bounce rays start this far along (and skip their own triangle) */
} rt_paths_params;

typedef struct rt_paths_out {
    float* radiance;          /* device, width*height float32, row-major */
    uint32_t* counts;         /* device, optional: bounces+1 uint32, rays traced at each depth */
} rt_paths_out;

/* Workspace: counters, per-sample radiance, two ray queues of width*height*spp 48-byte records. */
uint64_t rt_paths_workspace_bytes(uint32_t width, uint32_t height, uint32_t spp);
int rt_trace_paths(const rt_scene* scene, const rt_camera* cam, const rt_paths_params* params,
                   const rt_paths_out* out, void* workspace_dev, uint64_t workspace_bytes,
                   const rt_tuning* tuning, void* stream);

/* ---- mesh: rendering/_modeling.py:48-49 (Mesh.compute_normals, a stub in the reference) ----
 * Area-weighted per-vertex normals: n(v) = sum of cross(P1 - P0, P2 - P0) over the (triangle,
 * corner) pairs that use v, added corner 0 first, then 1, then 2, ascending triangle within each,
 * starting from +0; N(v) = n / max(sqrt((x*x + y*y) + z*z), 1e-300). All FP64 from the FP32
 * positions, rounded once to FP32: the result is a fixed function of the mesh. Vertices no triangle
 * uses get (0, 0, 0); out-of-range indices are clamped to n_vertices - 1 as in rt_scene_build.
 * N of vertex v is written as three floats at (char*)normals_out + v * out_stride: pass the N
 * field of the vertex records (e.g. vertices + 16, stride 80 for MeshVertex) to fill them in place,
 * or a dense float3 array (stride 12). The output must not overlap the positions. A non-indexed
 * mesh needs no workspace (NULL allowed). */
uint64_t rt_mesh_normals_workspace_bytes(uint32_t n_vertices, uint32_t n_tris);
int rt_mesh_compute_normals(const rt_mesh_desc* mesh, float* normals_out, uint32_t out_stride,
                            void* workspace_dev, uint64_t workspace_bytes, void* stream);

/* MeshVertex records (rendering/_modeling.py:22-28: 80 bytes, P@0, N@16, C@32, T@48, B@64, float3
 * fields padded to 16 bytes) from dense device arrays pos [n,3], nrm [n,3] and uv [n,2] floats;
 * nrm/uv may be NULL (written as zero), T, B and the pad words are zero. This is the device form of
 * filling `Mesh(vertices=...)` from arrays; `records` must be 16-byte aligned. */
int rt_mesh_pack(const float* pos, const float* nrm, const float* uv, uint32_t n_vertices,
                 void* records, void* stream);
/* The inverse: dense P / N / C arrays out of MeshVertex records; any output may be NULL. */
int rt_mesh_unpack(const void* records, uint32_t n_vertices, float* pos, float* nrm, float* uv,
                   void* stream);

/* ---- image export: rendering/_core.py:340 (RGBA render targets are BGRA UNORM_INT8 images) ----
 * n_pixels float RGB triples (e.g. rt_render's output) -> 4-byte B, G, R, A pixels with the OpenCL
 * UNORM_INT8 write conversion convert_uchar_sat_rte(f * 255) (NaN -> 0) and A = 255. */
int rt_image_to_bgra8(const float* rgb, uint32_t n_pixels, uint8_t* bgra, void* stream);

/* ---- textures: rendering/_core.py:551-578 (create_texture2D's default float4 texels) ----
 * n_texels float4 RGBA texels (16 B, the reference's texture format) -> RGBA8 texels (R, G, B, A
 * bytes) with the same UNORM_INT8 conversion, so a texture made by create_texture2D(w, h) can be
 * bound to rt_render once converted; `rgba` must be 16-byte aligned. */
int rt_texels_to_rgba8(const float* rgba, uint64_t n_texels, uint8_t* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RENDERTOY_RT_H */
