"""ORACLE / TEST INFRASTRUCTURE ONLY — never imported by the product (rendertoy_amd).

Python face of the CPU restatement in rt_twin.c (loaded through ctypes from oracle/librt_twin.so,
built by oracle/build.py) plus small numpy restatements used to pin it:

  morton_np          30-bit Morton codes, vectorised bit interleave (closed form)
  camera_dirs_fp64   primary-ray directions in FP64 from the look_at/perspective definitions of
                     rendering/_core.py:528-548 and the viewport of rendering/_raster.py:126-129
  bilinear_fp64      OpenCL 1.2 linear filter with repeat addressing, FP64

The reference ray caster is a stub (rendering/_raycaster.py:30-36) with no tests or golden data;
the pins are the closed forms, invariants and brute-force checks in tests/test_oracle_*.py.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librt_twin.so")
EMPTY_ROOT = 0x7FFFFFFF

_lib = None


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build as _b
            _b.build_oracle()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        i64 = ctypes.c_int64
        f32 = ctypes.c_float
        f64 = ctypes.c_double
        i32 = ctypes.c_int32
        L.tw_expand10.restype = ctypes.c_uint32
        L.tw_expand10.argtypes = [ctypes.c_uint32]
        L.tw_morton3.restype = ctypes.c_uint32
        L.tw_morton3.argtypes = [ctypes.c_uint32] * 3
        L.tw_build.restype = ctypes.c_int
        L.tw_build.argtypes = [i64] + [vp] * 6 + [vp] * 7 + [vp, ctypes.c_int]
        L.tw_trace_closest.restype = None
        L.tw_trace_closest.argtypes = [vp, vp, i32, i64] + [vp] * 8 + [f32, f32] + [vp] * 6
        L.tw_trace_any.restype = None
        L.tw_trace_any.argtypes = [vp, vp, i32, i64] + [vp] * 8 + [f32, f32, vp]
        L.tw_brute_closest32.restype = None
        L.tw_brute_closest32.argtypes = [vp, i64, i64] + [vp] * 6 + [f32, f32, vp, vp]
        L.tw_trace_camera.restype = None
        L.tw_trace_camera.argtypes = [vp, vp, i32, vp, ctypes.c_uint32, ctypes.c_uint32] + [vp] * 9 + [ctypes.c_int]
        L.tw_generate_rays.restype = None
        L.tw_generate_rays.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32] + [vp] * 6
        L.tw_render.restype = None
        L.tw_render.argtypes = [vp, vp, vp, i32, vp, ctypes.c_uint32, ctypes.c_uint32, vp,
                                ctypes.c_uint32] + [vp] * 9
        L.tw_philox4x32_10.restype = None
        L.tw_philox4x32_10.argtypes = [i64, vp, vp, vp]
        L.tw_cos_sin_2pi.restype = None
        L.tw_cos_sin_2pi.argtypes = [i64, vp, vp, vp]
        L.tw_cosine_dir.restype = None
        L.tw_cosine_dir.argtypes = [i64, vp, vp, vp, vp]
        L.tw_paths.restype = None
        L.tw_paths.argtypes = [vp, vp, i32, vp] + [ctypes.c_uint32] * 6 + [f32, f32, f32, vp, vp, vp]
        L.tw_closest64_brute.restype = None
        L.tw_closest64_brute.argtypes = [vp, i64, i64] + [vp] * 6 + [f64, f64, f64, f64] + [vp] * 5
        L.tw_closest64_bvh.restype = None
        L.tw_closest64_bvh.argtypes = [vp, i64, vp, vp, i32, i64] + [vp] * 6 + [f64, f64, f64, f64] + [vp] * 5
        L.tw_tex_sample.restype = None
        L.tw_tex_sample.argtypes = [vp, i32, i32, i64, vp, vp, i32, vp]
        L.tw_rgb_to_bgra8.restype = None
        L.tw_rgb_to_bgra8.argtypes = [vp, i64, vp]
        L.tw_raster_ids.restype = None
        L.tw_raster_ids.argtypes = [vp, i64, vp, ctypes.c_uint32, ctypes.c_uint32, f64, f64, vp, vp]
        L.tw_vertex_normals.restype = None
        L.tw_vertex_normals.argtypes = [vp, i64, vp, i64, vp]
        L.tw_num_threads.restype = ctypes.c_int
        L.tw_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def rgb_to_bgra8(rgb):
    """The twin of rt_image_to_bgra8: float RGB [..., 3] -> uint8 B, G, R, 255 [..., 4]."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    out = np.empty(rgb.shape[:-1] + (4,), np.uint8)
    lib().tw_rgb_to_bgra8(_p(rgb), rgb.size // 3, _p(out))
    return out


RASTER_EDGE, RASTER_DEPTH_TIE = 1, 2


def raster_ids(tri_pos, cam12, W, H, edge_eps=2e-3, depth_eps=1e-5):
    """Primary visibility by rasterisation (tw_raster_ids): the triangle each pixel centre sees by
    projection, near-plane clipping, screen-space edge functions and the nearest depth, in FP64.
    tri_pos [T,9] float32, cam12 the device camera (eye, right, up, forward). Returns prim [H*W]
    (-1 = background) and flags [H*W] (RASTER_EDGE: a triangle within edge_eps pixels of the centre
    lies at or in front of the winner; RASTER_DEPTH_TIE: a second covering triangle within
    depth_eps relative depth)."""
    tri_pos = np.ascontiguousarray(tri_pos, np.float32).reshape(-1, 9)
    cam12 = np.ascontiguousarray(cam12, np.float32)
    prim = np.empty(W * H, np.int32)
    flags = np.empty(W * H, np.uint8)
    lib().tw_raster_ids(_p(tri_pos), tri_pos.shape[0], _p(cam12), W, H, float(edge_eps),
                        float(depth_eps), _p(prim), _p(flags))
    return {"prim": prim, "flags": flags}


def vertex_normals(pos, idx=None):
    """Area-weighted per-vertex normals of the twin (tw_vertex_normals): pos [V,3] float32,
    idx [T,3] int32 or None (vertices 3t..3t+2, T = V // 3); returns [V,3] float32."""
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
    if idx is None:
        n_tris = pos.shape[0] // 3
    else:
        idx = np.ascontiguousarray(idx, dtype=np.int32).reshape(-1, 3)
        n_tris = idx.shape[0]
    out = np.empty_like(pos)
    lib().tw_vertex_normals(_p(pos), pos.shape[0], _p(idx), n_tris, _p(out))
    return out


def set_threads(n):
    lib().tw_set_threads(int(n))


def num_threads():
    return lib().tw_num_threads()


# ------------------------------------------------------------------------------------------
# numpy restatements (closed forms)
# ------------------------------------------------------------------------------------------

def expand10_np(v):
    """Spreads the low 10 bits of v to every third bit (bit k -> bit 3k), by definition."""
    v = np.asarray(v, dtype=np.uint64) & 0x3FF
    out = np.zeros_like(v)
    for k in range(10):
        out |= ((v >> np.uint64(k)) & np.uint64(1)) << np.uint64(3 * k)
    return out.astype(np.uint32)


def morton_np(x, y, z):
    return ((expand10_np(x).astype(np.uint64) << 2) | (expand10_np(y).astype(np.uint64) << 1)
            | expand10_np(z).astype(np.uint64)).astype(np.uint32)


def camera_dirs_fp64(eye, target, up, fov, W, H):
    """Primary directions [H*W, 3] in FP64: view-space (ndc.x/ws, ndc.y/hs, 1) with
    hs = 1/tan(fov/2), ws = hs/aspect, rotated by the look_at basis (rendering/_core.py:528-548),
    ndc from pixel centres with the viewport y flip (rendering/_raster.py:126-129, :298-299)."""
    eye, target, up = (np.asarray(a, dtype=np.float64) for a in (eye, target, up))
    z = target - eye
    z /= np.linalg.norm(z)
    x = np.cross(up, z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    hs = 1.0 / np.tan(fov / 2)
    ws = hs / (W / H)
    col, row = np.meshgrid(np.arange(W), np.arange(H))
    ndx = (col.ravel() + 0.5) / W * 2 - 1
    ndy = 1 - (row.ravel() + 0.5) / H * 2
    d = (ndx / ws)[:, None] * x + (ndy / hs)[:, None] * y + z
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def bilinear_fp64(tex, u, v):
    """OpenCL 1.2 sec. 8.2 linear filtering with CLK_ADDRESS_REPEAT on an RGBA8 image
    [H, W, 4] (normalised coordinates, row = v); returns RGB in [0, 1], FP64."""
    H, W = tex.shape[:2]
    u = np.asarray(u, np.float64)
    v = np.asarray(v, np.float64)
    x = (u - np.floor(u)) * W - 0.5
    y = (v - np.floor(v)) * H - 0.5
    i0 = np.floor(x).astype(np.int64)
    j0 = np.floor(y).astype(np.int64)
    a = x - i0
    b = y - j0
    i1 = (i0 + 1) % W
    j1 = (j0 + 1) % H
    i0 %= W
    j0 %= H
    t = tex[..., :3].astype(np.float64) / 255.0
    return (((1 - a) * (1 - b))[..., None] * t[j0, i0] + (a * (1 - b))[..., None] * t[j0, i1]
            + ((1 - a) * b)[..., None] * t[j1, i0] + (a * b)[..., None] * t[j1, i1])


# ------------------------------------------------------------------------------------------
# twin (FP32 restatement of the device algorithm)
# ------------------------------------------------------------------------------------------

class TwinScene:
    """The CPU twin's build of a scene from per-primitive arrays (scenes.triangle_arrays)."""

    BUILDERS = {"lbvh": 0, "ploc": 1}  # the device's RT_BUILD_* values

    def __init__(self, pos, nrm, uv, mesh_of, has_nrm, has_uv, builder="lbvh"):
        n = pos.shape[0]
        self.n = n
        self.pos = np.ascontiguousarray(pos, np.float32)
        self.keys = np.zeros(max(n, 1), np.uint32)
        self.order = np.zeros(max(n, 1), np.uint32)
        self.nodes = np.zeros((max(n - 1, 1), 16), np.float32)
        self.nodes4 = np.zeros((max(n - 1, 1), 32), np.float32)
        self.tris = np.zeros((max(n, 1), 12), np.float32)
        self.shade = np.zeros((max(n, 1), 16), np.float32)
        self.bounds = np.zeros(8, np.float32)
        root = ctypes.c_int32(0)
        hn = np.ascontiguousarray(has_nrm, np.uint8)
        hu = np.ascontiguousarray(has_uv, np.uint8)
        rc = lib().tw_build(n, _p(self.pos), _p(np.ascontiguousarray(nrm, np.float32)),
                            _p(np.ascontiguousarray(uv, np.float32)),
                            _p(np.ascontiguousarray(mesh_of, np.int32)), _p(hn), _p(hu),
                            _p(self.keys), _p(self.order), _p(self.nodes), _p(self.nodes4), _p(self.tris),
                            _p(self.shade), _p(self.bounds), ctypes.byref(root),
                            self.BUILDERS[builder])
        if rc != 0:
            raise MemoryError("tw_build failed")
        self.root = root.value
        self.keys, self.order = self.keys[:n], self.order[:n]
        self.tris, self.shade = self.tris[:n], self.shade[:n]
        self.nodes = self.nodes[:max(n - 1, 0)]
        self.nodes4 = self.nodes4[:max(n - 1, 0)]

    @property
    def nodes_full(self):
        return self.nodes if self.nodes.shape[0] else np.zeros((1, 16), np.float32)

    @property
    def nodes4_full(self):
        return self.nodes4 if self.nodes4.shape[0] else np.zeros((1, 32), np.float32)

    def trace_closest(self, o, d, tmin=0.0, tmax=np.inf, counters=False):
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        n = o.shape[0]
        cols = [np.ascontiguousarray(o[:, k]) for k in range(3)] + \
               [np.ascontiguousarray(d[:, k]) for k in range(3)]
        t = np.empty(n, np.float32)
        u = np.empty(n, np.float32)
        v = np.empty(n, np.float32)
        prim = np.empty(n, np.int32)
        slot = np.empty(n, np.int32)
        cnt = np.empty((n, 2), np.int32) if counters else None
        lib().tw_trace_closest(_p(self.nodes_full), _p(self.tris if self.n else np.zeros((1, 12), np.float32)),
                               self.root, n, *[_p(c) for c in cols], None, None, tmin, tmax,
                               _p(t), _p(u), _p(v), _p(prim), _p(slot), _p(cnt))
        out = {"t": t, "u": u, "v": v, "prim": prim, "slot": slot}
        if counters:
            out["node_visits"] = cnt[:, 0]
            out["tri_tests"] = cnt[:, 1]
        return out

    def trace_camera(self, cam12, W, H, quad=False):
        """Camera closest hit with the device's wave-coherent traversal over 8x8 tiles (quad: the
        quad-lane kernel's 16-ray packets over 4x4 tiles): dict of per-pixel t, u, v, prim, slot
        plus per-tile wave_nodes / wave_leaves / wave_tris (4-wide node records, leaves and
        triangle records fetched by the wave) and per-pixel lane_nodes / lane_tris (node records
        and triangles that pixel's own ray tested)."""
        n = W * H
        out = {"t": np.empty(n, np.float32), "u": np.empty(n, np.float32),
               "v": np.empty(n, np.float32), "prim": np.empty(n, np.int32),
               "slot": np.empty(n, np.int32)}
        ts = 4 if quad else 8
        ntiles = ((W + ts - 1) // ts) * ((H + ts - 1) // ts)
        out["wave_nodes"] = np.empty(ntiles, np.int64)
        out["wave_leaves"] = np.empty(ntiles, np.int64)
        out["wave_tris"] = np.empty(ntiles, np.int64)
        lane = np.empty((n, 2), np.int32)
        lib().tw_trace_camera(_p(self.nodes4_full), _p(self.tris if self.n else np.zeros((1, 12), np.float32)),
                              self.root, _p(np.ascontiguousarray(cam12, np.float32)), W, H,
                              _p(out["t"]), _p(out["u"]), _p(out["v"]), _p(out["prim"]),
                              _p(out["slot"]), _p(out["wave_nodes"]), _p(out["wave_leaves"]),
                              _p(out["wave_tris"]), _p(lane), 1 if quad else 0)
        out["lane_nodes"] = lane[:, 0]
        out["lane_tris"] = lane[:, 1]
        return out

    def trace_any(self, o, d, tmin=0.0, tmax=np.inf):
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        n = o.shape[0]
        cols = [np.ascontiguousarray(o[:, k]) for k in range(3)] + \
               [np.ascontiguousarray(d[:, k]) for k in range(3)]
        occ = np.empty(n, np.uint8)
        lib().tw_trace_any(_p(self.nodes_full), _p(self.tris if self.n else np.zeros((1, 12), np.float32)),
                           self.root, n, *[_p(c) for c in cols], None, None, tmin, tmax, _p(occ))
        return occ

    def brute_closest32(self, o, d, tmin=0.0, tmax=np.inf):
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        n = o.shape[0]
        cols = [np.ascontiguousarray(o[:, k]) for k in range(3)] + \
               [np.ascontiguousarray(d[:, k]) for k in range(3)]
        t = np.empty(n, np.float32)
        prim = np.empty(n, np.int32)
        lib().tw_brute_closest32(_p(self.tris), self.n, n, *[_p(c) for c in cols], tmin, tmax,
                                 _p(t), _p(prim))
        return t, prim

    def render(self, cam12, W, H, light, ambient=0.2, background=(0, 0, 0), shadow_tmin=1e-4,
               shadows=True, bilinear=True, textures=(), mesh_tex=None):
        params = np.array([*light, ambient, *background, shadow_tmin], np.float32)
        flags = (1 if shadows else 0) | (2 if bilinear else 0)
        rgb = np.empty((H, W, 3), np.float32)
        t = np.empty(H * W, np.float32)
        prim = np.empty(H * W, np.int32)
        stats = np.zeros(3, np.int64)
        if textures:
            blobs = [np.ascontiguousarray(x, np.uint8).reshape(-1) for x in textures]
            offs = np.cumsum([0] + [b.size for b in blobs[:-1]]).astype(np.int64)
            texcat = np.concatenate(blobs)
            tw = np.array([x.shape[1] for x in textures], np.int32)
            th = np.array([x.shape[0] for x in textures], np.int32)
        else:
            texcat, offs, tw, th = np.zeros(4, np.uint8), np.zeros(1, np.int64), np.ones(1, np.int32), np.ones(1, np.int32)
        mt = None if mesh_tex is None else np.ascontiguousarray(mesh_tex, np.int32)
        lib().tw_render(_p(self.nodes4_full), _p(self.tris if self.n else np.zeros((1, 12), np.float32)),
                        _p(self.shade if self.n else np.zeros((1, 16), np.float32)), self.root,
                        _p(np.ascontiguousarray(cam12, np.float32)), W, H, _p(params), flags,
                        _p(texcat), _p(offs), _p(tw), _p(th), _p(mt), _p(rgb), _p(t), _p(prim),
                        _p(stats))
        return rgb, t, prim, stats

    def paths(self, cam12, W, H, spp=4, bounces=2, seed=0x5EED, frame=0, albedo=0.7, sky=1.0,
              tmin=1e-4):
        """Diffuse paths of rt_trace_paths: (radiance [H*W], samples [H*W*spp], counts per depth)."""
        n = W * H
        rad = np.empty(n, np.float32)
        smp = np.empty(n * spp, np.float32)
        cnt = np.zeros(17, np.int64)
        lib().tw_paths(_p(self.nodes_full), _p(self.tris if self.n else np.zeros((1, 12), np.float32)),
                       self.root, _p(np.ascontiguousarray(cam12, np.float32)), W, H, spp, bounces,
                       seed, frame, albedo, sky, tmin, _p(smp), _p(rad), _p(cnt))
        return rad, smp, cnt[:bounces + 1]

    def closest64(self, o, d, tmin=0.0, edge_eps=1e-6, tie_eps=1e-5, margin=1e-4, brute=False):
        return closest64(self.pos, o, d, tmin, edge_eps, tie_eps, margin, None if brute else self)


def twin_tex_sample(tex, u, v, bilinear=True):
    """The device's RGBA8 texture fetch at (u, v) arrays: RGB float32 [n, 3]."""
    tex = np.ascontiguousarray(tex, np.uint8)
    u = np.ascontiguousarray(u, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    out = np.empty((u.size, 3), np.float32)
    lib().tw_tex_sample(_p(tex), tex.shape[1], tex.shape[0], u.size, _p(u), _p(v), int(bilinear), _p(out))
    return out


def twin_generate_rays(cam12, W, H):
    n = W * H
    cols = [np.empty(n, np.float32) for _ in range(6)]
    lib().tw_generate_rays(_p(np.ascontiguousarray(cam12, np.float32)), W, H, *[_p(c) for c in cols])
    return np.stack(cols[:3], 1), np.stack(cols[3:], 1)


EDGE, DEPTH_TIE, NEAR_MISS = 1, 2, 4


def closest64(tri_pos, o, d, tmin=0.0, edge_eps=1e-6, tie_eps=1e-5, margin=1e-4, scene=None):
    """FP64 closest hit over the triangles whose FP64 box, inflated by `margin`, the ray crosses
    (brute force over all boxes, or found through `scene`'s hierarchy: the same set): dict t (float64, inf on miss), prim (-1 on miss), u, v, and `edge`, a bit set of
    EDGE (winner within max(edge_eps, FP32 error bound) barycentric of an edge), DEPTH_TIE
    (another hit within max(tie_eps relative, FP32 t error bounds) of the winner), NEAR_MISS (a
    triangle missed by less than its FP32 barycentric error bound within that range): the rays on
    which an FP32 Moller-Trumbore evaluation may legitimately differ; 0 elsewhere."""
    tri_pos = np.ascontiguousarray(tri_pos, np.float32)
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    n = o.shape[0]
    cols = [np.ascontiguousarray(o[:, k]) for k in range(3)] + \
           [np.ascontiguousarray(d[:, k]) for k in range(3)]
    t = np.empty(n, np.float64)
    prim = np.empty(n, np.int32)
    u = np.empty(n, np.float64)
    v = np.empty(n, np.float64)
    edge = np.empty(n, np.uint8)
    if scene is None:
        lib().tw_closest64_brute(_p(tri_pos), tri_pos.shape[0], n, *[_p(c) for c in cols], tmin,
                                 edge_eps, tie_eps, margin, _p(t), _p(prim), _p(u), _p(v), _p(edge))
    else:
        order = np.ascontiguousarray(scene.order if scene.n else np.zeros(1, np.uint32), np.uint32)
        lib().tw_closest64_bvh(_p(tri_pos), tri_pos.shape[0], _p(scene.nodes_full), _p(order),
                               scene.root, n, *[_p(c) for c in cols], tmin, edge_eps, tie_eps,
                               margin, _p(t), _p(prim), _p(u), _p(v), _p(edge))
    return {"t": t, "prim": prim, "u": u, "v": v, "edge": edge}


def shade64(tri_pos, tri_nrm, tri_uv, mesh_of, has_nrm, o, d, light, ambient=0.2,
            background=(0.0, 0.0, 0.0), shadow_tmin=1e-4, shadows=True, textures=(),
            mesh_tex=None, scene=None, primary=None):
    """FP64 shading of rays o + t d (the device's shading formula evaluated in FP64 on the FP64
    closest hit): rgb [n, 3] float64 and flags [n] (bit set of the primary ray's FP64 flags,
    plus 8 when the shadow ray is ambiguous). rgb = albedo * (ambient + V * max(0, N.L)), N the
    normalised barycentric interpolation (w0, u, v) of the vertex normals (geometric normal for
    meshes without normals), albedo the OpenCL linear filter of the mesh's RGBA8 texture (1 when
    none), V = 0 when the shadow ray from the hit point toward L hits anything beyond shadow_tmin.
    `primary`: the closest64 result of the same rays, if already computed."""
    tri_pos = np.ascontiguousarray(tri_pos, np.float32)
    prim = primary if primary is not None else closest64(tri_pos, o, d, scene=scene)
    n = len(o)
    rgb = np.tile(np.asarray(background, np.float64), (n, 1))
    flags = prim["edge"].astype(np.int32)
    hit = prim["prim"] >= 0
    p = prim["prim"][hit]
    u = prim["u"][hit]
    v = prim["v"][hit]
    w0 = 1.0 - u - v
    P = tri_pos[p].astype(np.float64).reshape(-1, 3, 3)
    Nv = np.asarray(tri_nrm, np.float32)[p].astype(np.float64).reshape(-1, 3, 3)
    geo = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    geo /= np.maximum(np.linalg.norm(geo, axis=1, keepdims=True), 1e-300)
    mh = np.asarray(has_nrm, bool)[np.asarray(mesh_of)[p]]
    N = w0[:, None] * Nv[:, 0] + u[:, None] * Nv[:, 1] + v[:, None] * Nv[:, 2]
    N = np.where(mh[:, None], N, geo)
    ln = np.linalg.norm(N, axis=1, keepdims=True)
    N = np.where(ln > 0, N / np.where(ln > 0, ln, 1), 0)
    L = np.asarray(light, np.float64)
    ndl = np.maximum(N @ L, 0.0)
    vis = np.ones(len(p))
    if shadows:
        lit = ndl > 0
        hp = o[hit].astype(np.float64) + prim["t"][hit][:, None] * d[hit].astype(np.float64)
        so = hp[lit].astype(np.float32)
        sd = np.tile(np.asarray(light, np.float32), (int(lit.sum()), 1))
        sh = closest64(tri_pos, so, sd, tmin=shadow_tmin, scene=scene)
        occ = sh["prim"] >= 0
        vis_l = np.where(occ, 0.0, 1.0)
        vis[lit] = vis_l
        sflag = np.zeros(len(p), np.int32)
        sflag[lit] = np.where(sh["edge"] != 0, 8, 0)
        # a shadow ray starting within the start offset of a surface is ambiguous too
        near = np.zeros(len(p), bool)
        near[lit] = np.abs(sh["t"] - shadow_tmin) < 1e-5
        flags[np.nonzero(hit)[0]] |= sflag | np.where(near, 8, 0)
    alb = np.ones((len(p), 3))
    if textures and mesh_tex is not None:
        UV = np.asarray(tri_uv, np.float32)[p].astype(np.float64).reshape(-1, 3, 2)
        tu = w0 * UV[:, 0, 0] + u * UV[:, 1, 0] + v * UV[:, 2, 0]
        tv = w0 * UV[:, 0, 1] + u * UV[:, 1, 1] + v * UV[:, 2, 1]
        ti = np.asarray(mesh_tex)[np.asarray(mesh_of)[p]]
        for k, tex in enumerate(textures):
            sel = ti == k
            if sel.any():
                alb[sel] = bilinear_fp64(np.asarray(tex), tu[sel], tv[sel])
    rgb[hit] = alb * (ambient + vis * ndl)[:, None]
    return rgb, flags


# ---- Philox4x32-10 and the path sampling (rt_trace_paths) ----

def twin_philox(ctr, key):
    """Philox4x32-10 of the C restatement: ctr [n, 4] uint32, key (k0, k1) -> [n, 4] uint32."""
    ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, 4)
    out = np.empty_like(ctr)
    k = np.ascontiguousarray(key, np.uint32)
    lib().tw_philox4x32_10(len(ctr), _p(ctr), _p(k), _p(out))
    return out


def philox_np(ctr, key):
    """Philox4x32-10 vectorised in numpy (uint64 products), from the SC'11 definition: round
    (c0..c3) -> (hi(M1 c2) ^ c1 ^ k0, lo(M1 c2), hi(M0 c0) ^ c3 ^ k1, lo(M0 c0)), M0 = 0xD2511F53,
    M1 = 0xCD9E8D57; key += (0x9E3779B9, 0xBB67AE85) between rounds; 10 rounds."""
    c = np.array(ctr, dtype=np.uint64).reshape(-1, 4)
    k0 = np.uint64(key[0])
    k1 = np.uint64(key[1])
    m32 = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = np.uint64(0xD2511F53) * c[:, 0]
        p1 = np.uint64(0xCD9E8D57) * c[:, 2]
        c = np.stack([(p1 >> np.uint64(32)) ^ c[:, 1] ^ k0, p1 & m32,
                      (p0 >> np.uint64(32)) ^ c[:, 3] ^ k1, p0 & m32], axis=1)
        k0 = (k0 + np.uint64(0x9E3779B9)) & m32
        k1 = (k1 + np.uint64(0xBB67AE85)) & m32
    return c.astype(np.uint32)


def twin_cos_sin_2pi(x):
    x = np.ascontiguousarray(x, np.float32)
    c = np.empty_like(x)
    s = np.empty_like(x)
    lib().tw_cos_sin_2pi(len(x), _p(x), _p(c), _p(s))
    return c, s


def twin_cosine_dir(n, u0, u1):
    n = np.ascontiguousarray(n, np.float32).reshape(-1, 3)
    u0 = np.ascontiguousarray(u0, np.float32)
    u1 = np.ascontiguousarray(u1, np.float32)
    d = np.empty_like(n)
    lib().tw_cosine_dir(len(n), _p(n), _p(u0), _p(u1), _p(d))
    return d
