// rt_mesh.hip — mesh data on the device. Per-vertex normals: Mesh.compute_normals
// (rendering/_modeling.py:48-49, a stub in the reference that raises "Not implemented yet").
//
// Definition (area-weighted vertex normals):
//   n(v) = sum of cross(P1 - P0, P2 - P0) over the (triangle t, corner k) pairs with idx[t][k] == v,
//          added in the order k = 0, 1, 2 and, within k, ascending t, starting from +0;
//   N(v) = n(v) / d, d = sqrt((x*x + y*y) + z*z) raised to 1e-300 (a NaN d stays NaN).
// Everything is FP64 from the FP32 positions, rounded once to FP32, so N is a fixed function of the
// mesh: the sum order is fixed by a stable sort of the pairs by vertex, not by scheduling.
// Vertices that no triangle uses get (0, 0, 0). Out-of-range indices are clamped to n_vertices - 1,
// as the scene build does.
//
// Pipeline (one stream):
//   1. k_nrm_face   per triangle: FP64 face normal (32 B record) and its 3 pairs
//                   (key = vertex, value = p = k * n_tris + t, stored at position p).
//   2. sort         stable LSD radix sort of the pairs by vertex (ceil(bits(n_vertices - 1) / 8)
//                   passes), so each vertex's pairs form a run in ascending p.
//   3. k_nrm_runs   marks the first position of each run in run_of[vertex] (cleared beforehand).
//   4. k_nrm_vertex per vertex: sums its run, normalises, writes float3 N at out + v * out_stride.
// A non-indexed mesh (vertices 3t, 3t+1, 3t+2) has one pair per vertex: k_nrm_direct computes it
// with no sort and no workspace.
//
// Also here: MeshVertex records <-> dense float arrays (k_mesh_pack / k_mesh_unpack), the device-side
// ingest and export of the 80-byte vertex layout of rendering/_modeling.py:22-28.
#include <cmath>

#include "rt_common.h"
#include "rt_sort.h"

namespace rt {
namespace {

constexpr uint32_t kNoRun = 0xFFFFFFFFu;

struct NrmWorkspace {
    double4* face;    // n_tris
    uint32_t* keys0;  // 3 n_tris
    uint32_t* vals0;
    uint32_t* keys1;
    uint32_t* vals1;
    uint32_t* hist;
    uint32_t* totals;  // kRadix
    uint32_t* run_of;  // n_vertices
    uint64_t bytes;
};

inline uint64_t align256n(uint64_t x) { return (x + 255) & ~uint64_t(255); }

NrmWorkspace carve_nrm(char* base, uint64_t n_vertices, uint64_t n_tris) {
    NrmWorkspace w{};
    const uint64_t np = 3 * n_tris;
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        char* p = base ? base + off : nullptr;
        off = align256n(off + (bytes ? bytes : 4));
        return p;
    };
    w.face = (double4*)take(32 * n_tris);
    w.keys0 = (uint32_t*)take(4 * np);
    w.vals0 = (uint32_t*)take(4 * np);
    w.keys1 = (uint32_t*)take(4 * np);
    w.vals1 = (uint32_t*)take(4 * np);
    w.hist = (uint32_t*)take(sort_hist_bytes(np));
    w.totals = (uint32_t*)take(4 * kRadix);
    w.run_of = (uint32_t*)take(4 * n_vertices);
    w.bytes = off;
    return w;
}

RT_DEV double3 load_p64(const char* vb, uint32_t i, uint32_t stride, uint32_t off) {
    const float* p = (const float*)(vb + (uint64_t)i * stride + off);
    return make_double3((double)p[0], (double)p[1], (double)p[2]);
}

RT_DEV uint32_t clamp_index(int32_t i, uint32_t nv) {
    const uint32_t u = (uint32_t)i;
    return u < nv ? u : nv - 1;
}

// cross(b - a, c - a) in FP64 (no contraction: the library is built with -ffp-contract=off)
RT_DEV double3 face_normal(const double3 a, const double3 b, const double3 c) {
    const double e1x = b.x - a.x, e1y = b.y - a.y, e1z = b.z - a.z;
    const double e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
    return make_double3(e1y * e2z - e1z * e2y, e1z * e2x - e1x * e2z, e1x * e2y - e1y * e2x);
}

RT_DEV void store_normal(char* out, uint32_t v, uint32_t stride, double x, double y, double z) {
    const double sq = (x * x + y * y) + z * z;
    double d = sqrt(sq);
    d = d < 1e-300 ? 1e-300 : d;  // a NaN d is kept
    float* o = (float*)(out + (uint64_t)v * stride);
    o[0] = (float)(x / d);
    o[1] = (float)(y / d);
    o[2] = (float)(z / d);
}

__global__ void __launch_bounds__(256) k_nrm_face(rt_mesh_desc m, double4* __restrict__ face,
                                                   uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m.n_tris) return;
    const char* vb = (const char*)m.vertices;
    const uint32_t i0 = clamp_index(m.indices[3 * (uint64_t)t], m.n_vertices);
    const uint32_t i1 = clamp_index(m.indices[3 * (uint64_t)t + 1], m.n_vertices);
    const uint32_t i2 = clamp_index(m.indices[3 * (uint64_t)t + 2], m.n_vertices);
    const double3 f = face_normal(load_p64(vb, i0, m.vertex_stride, m.pos_offset),
                                  load_p64(vb, i1, m.vertex_stride, m.pos_offset),
                                  load_p64(vb, i2, m.vertex_stride, m.pos_offset));
    face[t] = make_double4(f.x, f.y, f.z, 0.0);
    const uint32_t T = m.n_tris;
    keys[t] = i0;
    keys[T + t] = i1;
    keys[2 * T + t] = i2;
    vals[t] = t;
    vals[T + t] = T + t;
    vals[2 * T + t] = 2 * T + t;
}

__global__ void __launch_bounds__(256) k_nrm_clear(uint32_t* __restrict__ run_of, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) run_of[i] = kNoRun;
}

// run_of[key] = first sorted position holding that key
__global__ void __launch_bounds__(256) k_nrm_runs(const uint32_t* __restrict__ keys, uint32_t np,
                                                   uint32_t* __restrict__ run_of) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) run_of[k] = i;
}

__global__ void __launch_bounds__(256) k_nrm_vertex(const uint32_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ vals, uint32_t np,
                                                     uint32_t n_tris, const double4* __restrict__ face,
                                                     const uint32_t* __restrict__ run_of,
                                                     uint32_t n_vertices, char* __restrict__ out,
                                                     uint32_t out_stride) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_vertices) return;
    double x = 0.0, y = 0.0, z = 0.0;
    const uint32_t first = run_of[v];
    if (first == kNoRun) {
        float* o = (float*)(out + (uint64_t)v * out_stride);
        o[0] = 0.0f;
        o[1] = 0.0f;
        o[2] = 0.0f;
        return;
    }
    for (uint32_t j = first; j < np && keys[j] == v; ++j) {
        const uint32_t p = vals[j];
        const uint32_t t = p >= 2 * n_tris ? p - 2 * n_tris : (p >= n_tris ? p - n_tris : p);
        const double4 f = face[t];
        x = x + f.x;
        y = y + f.y;
        z = z + f.z;
    }
    store_normal(out, v, out_stride, x, y, z);
}

// Non-indexed mesh: vertex v < 3 n_tris is corner v % 3 of triangle v / 3, its only pair.
__global__ void __launch_bounds__(256) k_nrm_direct(rt_mesh_desc m, char* __restrict__ out,
                                                     uint32_t out_stride) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m.n_vertices) return;
    if (v >= 3 * (uint64_t)m.n_tris) {
        float* o = (float*)(out + (uint64_t)v * out_stride);
        o[0] = 0.0f;
        o[1] = 0.0f;
        o[2] = 0.0f;
        return;
    }
    const uint32_t t0 = v - v % 3u;
    const char* vb = (const char*)m.vertices;
    const double3 f = face_normal(load_p64(vb, t0, m.vertex_stride, m.pos_offset),
                                  load_p64(vb, t0 + 1, m.vertex_stride, m.pos_offset),
                                  load_p64(vb, t0 + 2, m.vertex_stride, m.pos_offset));
    store_normal(out, v, out_stride, 0.0 + f.x, 0.0 + f.y, 0.0 + f.z);
}

// MeshVertex records (80 B: P@0, N@16, C@32, T@48, B@64, float3 fields padded to 16 B) from
// dense float arrays; absent attributes, T, B and the pad words are written as zero. A wave writes
// the 64 records of its vertices as 320 consecutive 16-byte chunks, lane l taking chunks l + 64j,
// so each store instruction covers 1 KiB of contiguous output; chunk c is field c % 5 of vertex
// c / 5.
__global__ void __launch_bounds__(256) k_mesh_pack(const float* __restrict__ pos,
                                                    const float* __restrict__ nrm,
                                                    const float* __restrict__ uv, uint32_t n,
                                                    float4* __restrict__ rec) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t v0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u);  // the wave's first vertex
    if (v0 >= n) return;
    const uint32_t nw = n - v0 < 64 ? (uint32_t)(n - v0) : 64u;
#pragma unroll
    for (uint32_t j = 0; j < 5; ++j) {
        const uint32_t c = lane + 64u * j, vl = c / 5u, f = c - 5u * vl;
        if (vl >= nw) break;
        const uint64_t v = v0 + vl;
        float4 x = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (f == 0) {
            x = make_float4(pos[3 * v], pos[3 * v + 1], pos[3 * v + 2], 0.0f);
        } else if (f == 1) {
            if (nrm) x = make_float4(nrm[3 * v], nrm[3 * v + 1], nrm[3 * v + 2], 0.0f);
        } else if (f == 2) {
            if (uv) x = make_float4(uv[2 * v], uv[2 * v + 1], 0.0f, 0.0f);
        }
        rec[5 * v0 + c] = x;
    }
}

// Dense P / N / C arrays out of MeshVertex records (any output may be NULL).
__global__ void __launch_bounds__(256) k_mesh_unpack(const float4* __restrict__ rec, uint32_t n,
                                                      float* __restrict__ pos,
                                                      float* __restrict__ nrm,
                                                      float* __restrict__ uv) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const uint64_t i3 = 3 * (uint64_t)v, i2 = 2 * (uint64_t)v;
    const float4* r = rec + 5 * (uint64_t)v;
    if (pos) {
        const float4 a = r[0];
        pos[i3] = a.x;
        pos[i3 + 1] = a.y;
        pos[i3 + 2] = a.z;
    }
    if (nrm) {
        const float4 b = r[1];
        nrm[i3] = b.x;
        nrm[i3 + 1] = b.y;
        nrm[i3 + 2] = b.z;
    }
    if (uv) {
        const float4 c = r[2];
        uv[i2] = c.x;
        uv[i2 + 1] = c.y;
    }
}

inline uint32_t nblocks(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" uint64_t rt_mesh_normals_workspace_bytes(uint32_t n_vertices, uint32_t n_tris) {
    return carve_nrm(nullptr, n_vertices, n_tris).bytes;
}

extern "C" int rt_mesh_compute_normals(const rt_mesh_desc* mesh, float* normals_out,
                                       uint32_t out_stride, void* workspace, uint64_t ws_bytes,
                                       void* stream) {
    if (!mesh) return set_error(RT_EINVAL, "rt_mesh_compute_normals: mesh is NULL");
    const rt_mesh_desc& m = *mesh;
    if (m.n_vertices == 0) {
        if (m.n_tris) return set_error(RT_EINVAL, "rt_mesh_compute_normals: mesh has no vertices");
        return RT_OK;
    }
    if (!m.vertices) return set_error(RT_EINVAL, "rt_mesh_compute_normals: mesh vertices NULL");
    if (m.vertex_stride < 12 || (m.vertex_stride & 3) || (m.pos_offset & 3) ||
        m.pos_offset + 12 > m.vertex_stride)
        return set_error(RT_EINVAL, "rt_mesh_compute_normals: bad vertex stride/position offset");
    if (!m.indices && (uint64_t)m.n_tris * 3 > m.n_vertices)
        return set_error(RT_EINVAL, "rt_mesh_compute_normals: non-indexed mesh needs 3 vertices per triangle");
    if ((uint64_t)m.n_tris * 3 > 0xFFFFFFFFull)
        return set_error(RT_ECAPACITY, "rt_mesh_compute_normals: too many triangles");
    if (!normals_out) return set_error(RT_EINVAL, "rt_mesh_compute_normals: normals_out is NULL");
    if (out_stride < 12 || (out_stride & 3) || ((uintptr_t)normals_out & 3))
        return set_error(RT_EINVAL, "rt_mesh_compute_normals: output must be 4-byte aligned with stride >= 12");
    hipStream_t s = (hipStream_t)stream;
    char* out = (char*)normals_out;
    if (!m.indices) {
        hipLaunchKernelGGL(k_nrm_direct, dim3(nblocks(m.n_vertices, 256)), dim3(256), 0, s, m, out,
                           out_stride);
        return check_launch("rt_mesh_compute_normals");
    }
    const NrmWorkspace need = carve_nrm(nullptr, m.n_vertices, m.n_tris);
    if (!workspace || ws_bytes < need.bytes)
        return set_error(RT_EINVAL, "rt_mesh_compute_normals: workspace too small (need " +
                                        std::to_string(need.bytes) + " bytes)");
    if ((uintptr_t)workspace & 255)
        return set_error(RT_EINVAL, "rt_mesh_compute_normals: workspace must be 256-byte aligned");
    NrmWorkspace w = carve_nrm((char*)workspace, m.n_vertices, m.n_tris);
    const uint32_t np = 3 * m.n_tris;
    hipLaunchKernelGGL(k_nrm_clear, dim3(nblocks(m.n_vertices, 256)), dim3(256), 0, s, w.run_of,
                       m.n_vertices);
    uint32_t *ki = w.keys0, *vi = w.vals0;
    if (np) {
        hipLaunchKernelGGL(k_nrm_face, dim3(nblocks(m.n_tris, 256)), dim3(256), 0, s, m, w.face, ki, vi);
        const int bits = m.n_vertices > 1 ? 32 - __builtin_clz(m.n_vertices - 1) : 0;
        sort_pairs(ki, vi, w.keys1, w.vals1, np, bits, w.hist, w.totals, s);
        hipLaunchKernelGGL(k_nrm_runs, dim3(nblocks(np, 256)), dim3(256), 0, s, ki, np, w.run_of);
    }
    hipLaunchKernelGGL(k_nrm_vertex, dim3(nblocks(m.n_vertices, 256)), dim3(256), 0, s, ki, vi, np,
                       m.n_tris, w.face, w.run_of, m.n_vertices, out, out_stride);
    return check_launch("rt_mesh_compute_normals");
}

extern "C" int rt_mesh_pack(const float* pos, const float* nrm, const float* uv, uint32_t n_vertices,
                            void* records, void* stream) {
    if (n_vertices == 0) return RT_OK;
    if (!pos || !records) return set_error(RT_EINVAL, "rt_mesh_pack: positions and records are required");
    if (((uintptr_t)pos | (uintptr_t)nrm | (uintptr_t)uv) & 3)
        return set_error(RT_EINVAL, "rt_mesh_pack: arrays must be 4-byte aligned");
    if ((uintptr_t)records & 15) return set_error(RT_EINVAL, "rt_mesh_pack: records must be 16-byte aligned");
    hipLaunchKernelGGL(k_mesh_pack, dim3(nblocks(n_vertices, 256)), dim3(256), 0, (hipStream_t)stream,
                       pos, nrm, uv, n_vertices, (float4*)records);
    return check_launch("rt_mesh_pack");
}

extern "C" int rt_mesh_unpack(const void* records, uint32_t n_vertices, float* pos, float* nrm, float* uv,
                              void* stream) {
    if (n_vertices == 0) return RT_OK;
    if (!records) return set_error(RT_EINVAL, "rt_mesh_unpack: records is NULL");
    if ((uintptr_t)records & 15) return set_error(RT_EINVAL, "rt_mesh_unpack: records must be 16-byte aligned");
    if (((uintptr_t)pos | (uintptr_t)nrm | (uintptr_t)uv) & 3)
        return set_error(RT_EINVAL, "rt_mesh_unpack: arrays must be 4-byte aligned");
    hipLaunchKernelGGL(k_mesh_unpack, dim3(nblocks(n_vertices, 256)), dim3(256), 0, (hipStream_t)stream,
                       (const float4*)records, n_vertices, pos, nrm, uv);
    return check_launch("rt_mesh_unpack");
}
