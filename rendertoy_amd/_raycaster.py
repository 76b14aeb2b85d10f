"""The ray caster of rendering/_raycaster.py on the MI355X kernels of librtx_hip.so.

Reference API kept: `Raycaster(models: List[Mesh])` builds the acceleration structure
(`_build_ads`, rendering/_raycaster.py:26-33) and `ray_cast(rays) -> array` answers closest-hit
queries (:35-36); the `BVH_AABB` and `BVH_Triangle` structs (:8-22) are registered with the same
layout. Additions: `Ray`/`Hit` structs, `occluded(rays)`, `render(...)` and `generate_rays(...)`.
Triangle identity of a hit is (mesh, index) as BVH_Triangle defines it (:16-22).
Every call runs on the GPU; there is no CPU path.
"""
import ctypes
import math
import typing

import numpy as np
import torch

from . import _native as N
from ._core import (kernel_struct, float3, float4, DeviceArray, create_buffer, Dispatcher,
                    to_array, texture_pointer, RGBA)
from ._modeling import Mesh, MeshVertex


@kernel_struct
class BVH_AABB:
    min: float3
    max: float3
    count: int
    elements: int


@kernel_struct
class BVH_Triangle:
    v0: float3
    v1: float3
    v2: float3
    index: int
    mesh: int


@kernel_struct
class Ray:
    """origin @0, direction @16 (32 bytes)."""
    origin: float3
    direction: float3


@kernel_struct
class Hit:
    """t, barycentrics (u, v) of v1 and v2, triangle index within its mesh, mesh position in
    `models`; a miss is t=+inf, index=mesh=-1 (20 bytes)."""
    t: np.float32
    u: np.float32
    v: np.float32
    index: np.int32
    mesh: np.int32


class Camera:
    """Pinhole camera of look_at + perspective (rendering/_core.py:528-548) reduced to the vectors
    the kernels use. Host math in float32 as the reference's make_float* values are."""

    def __init__(self, eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=math.pi / 4,
                 aspect=1.0):
        eye = np.asarray(to_array(eye) if np.asarray(eye).dtype.names else eye, dtype=np.float32)
        target = np.asarray(to_array(target) if np.asarray(target).dtype.names else target,
                            dtype=np.float32)
        up = np.asarray(to_array(up) if np.asarray(up).dtype.names else up, dtype=np.float32)
        z = target - eye
        z = (z / np.float32(np.sqrt(np.dot(z, z)))).astype(np.float32)
        x = np.cross(up, z).astype(np.float32)
        x = (x / np.float32(np.sqrt(np.dot(x, x)))).astype(np.float32)
        y = np.cross(z, x).astype(np.float32)
        tan_half = math.tan(fov / 2.0)
        self.eye = eye.astype(np.float32)
        self.right = (x * np.float32(aspect * tan_half)).astype(np.float32)
        self.up = (y * np.float32(tan_half)).astype(np.float32)
        self.forward = z
        self.fov = float(fov)
        self.aspect = float(aspect)

    def as_array(self):
        """12 float32: eye, right, up, forward."""
        return np.concatenate([self.eye, self.right, self.up, self.forward]).astype(np.float32)

    def to_c(self):
        c = N.rt_camera()
        for name in ("eye", "right", "up", "forward"):
            getattr(c, name)[:] = [float(v) for v in getattr(self, name)]
        return c


def _ptr(t) -> int:
    if t is None:
        return None
    if isinstance(t, DeviceArray):
        return t.data_ptr()
    return t.data_ptr()


def _check_device(t, what):
    dev = t.device if not isinstance(t, DeviceArray) else t.storage.device
    if dev.type != "cuda":
        raise RuntimeError(f"{what} must live on the HIP device (got {dev})")


def _fill_mesh_desc(d, m, keep):
    """Fills rt_mesh_desc `d` from Mesh `m` (MeshVertex records, optional index buffer); narrowed
    index copies are appended to `keep`, which must outlive the call that reads them."""
    vb = m.vertices
    _check_device(vb, "Mesh.vertices")
    if vb.dtype != MeshVertex:
        raise TypeError("Mesh.vertices must be a buffer of MeshVertex")
    d.vertices = vb.data_ptr()
    d.n_vertices = vb.shape[0]
    d.n_tris = vb.shape[0] // 3 if m.indices is None else m.indices.size // 3  # :33
    d.vertex_stride = MeshVertex.itemsize
    d.pos_offset = MeshVertex.fields["P"][1]
    d.nrm_offset = MeshVertex.fields["N"][1] if getattr(m, "has_normals", True) else N.RT_ABSENT
    d.uv_offset = MeshVertex.fields["C"][1] if getattr(m, "has_uvs", True) else N.RT_ABSENT
    if m.indices is not None:
        ib = m.indices
        _check_device(ib, "Mesh.indices")
        if ib.dtype != np.int32:
            # load_obj-style int64 index buffers (rendering/_loaders.py:17) are narrowed
            t = ib.tensor(torch.int64 if ib.dtype == np.int64 else None)
            ib32 = t.to(torch.int32).contiguous()
            keep.append(ib32)
            d.indices = ib32.data_ptr()
        else:
            d.indices = ib.data_ptr()
    else:
        d.indices = None


def compute_normals(mesh: Mesh, out: torch.Tensor = None, stream=None):
    """Area-weighted per-vertex normals of `mesh` on the device (Mesh.compute_normals,
    rendering/_modeling.py:48-49, is a stub in the reference).

    N(v) = n / |n| with n the FP64 sum of cross(P1 - P0, P2 - P0) over the triangles using v
    (corner-0 pairs first, then corner 1, then corner 2, ascending triangle within each), rounded
    once to float32; vertices no triangle uses get (0, 0, 0). With `out` = None the normals are
    written into the N field of the mesh's vertex records (and the mesh is marked as having
    normals); otherwise into `out`, a contiguous float32 device tensor [n_vertices, 3].
    Asynchronous on `stream` (default: the current stream)."""
    lib = N.require_gpu()
    d = N.rt_mesh_desc()
    keep = []
    # every allocation below is made on the stream the kernels run on, so torch's allocator
    # orders any reuse of the freed index copy and workspace after them
    ts = stream if stream is not None else torch.cuda.current_stream()
    with torch.cuda.stream(ts):
        _fill_mesh_desc(d, mesh, keep)
        if out is None:
            dst, stride = mesh.vertices.data_ptr() + MeshVertex.fields["N"][1], MeshVertex.itemsize
        else:
            _check_device(out, "out")
            if (out.dtype != torch.float32 or not out.is_contiguous()
                    or tuple(out.shape) != (d.n_vertices, 3)):
                raise TypeError(f"out must be a contiguous float32 tensor of shape ({d.n_vertices}, 3)")
            dst, stride = out.data_ptr(), 12
        ws, nbytes = None, 0
        if mesh.indices is not None:
            nbytes = lib.rt_mesh_normals_workspace_bytes(d.n_vertices, d.n_tris)
            ws = torch.empty(nbytes, dtype=torch.uint8, device=torch.device("cuda"))
        N.check(lib.rt_mesh_compute_normals(ctypes.byref(d), dst, stride,
                                            None if ws is None else ws.data_ptr(), nbytes,
                                            N.stream_handle(ts)), "rt_mesh_compute_normals")
        del ws, keep
    if out is None:
        mesh.has_normals = True
    return out


def _dense(t, cols, what, n=None):
    if t is None:
        return None
    _check_device(t, what)
    if t.dtype != torch.float32 or not t.is_contiguous() or t.dim() != 2 or t.shape[1] != cols:
        raise TypeError(f"{what} must be a contiguous float32 device tensor [n, {cols}]")
    if n is not None and t.shape[0] != n:
        raise ValueError(f"{what} has {t.shape[0]} rows, expected {n}")
    return t


def mesh_from_tensors(positions: torch.Tensor, indices: torch.Tensor = None, normals: torch.Tensor = None,
                      uvs: torch.Tensor = None, stream=None) -> Mesh:
    """A Mesh built on the device from device tensors (positions [V,3], optional normals [V,3] and
    uvs [V,2] float32, indices [T,3] int32/int64): the MeshVertex records (rendering/_modeling.py:
    22-28) are written by rt_mesh_pack, with no host round trip. The device counterpart of
    mesh_from_arrays; indices are kept as an int32 device buffer."""
    lib = N.require_gpu()
    pos = _dense(positions, 3, "positions")
    n = pos.shape[0]
    nrm = _dense(normals, 3, "normals", n)
    uv = _dense(uvs, 2, "uvs", n)
    ts = stream if stream is not None else torch.cuda.current_stream()
    with torch.cuda.stream(ts):
        vb = DeviceArray(torch.empty(max(n * MeshVertex.itemsize, 1), dtype=torch.uint8,
                                     device=torch.device("cuda")), MeshVertex, (n,))
        N.check(lib.rt_mesh_pack(pos.data_ptr(), None if nrm is None else nrm.data_ptr(),
                                 None if uv is None else uv.data_ptr(), n, vb.data_ptr(),
                                 N.stream_handle(ts)), "rt_mesh_pack")
        ib = None
        if indices is not None:
            _check_device(indices, "indices")
            if indices.dtype not in (torch.int32, torch.int64):
                raise TypeError("indices must be an int32 or int64 device tensor")
            if indices.dtype == torch.int64 and indices.numel() and (
                    int(indices.max()) >= 2 ** 31 or int(indices.min()) < -2 ** 31):
                raise ValueError("indices must fit int32")
            i32 = indices.reshape(-1).to(torch.int32).contiguous()
            ib = DeviceArray(i32.view(torch.uint8), np.int32, (i32.numel(),))
    m = Mesh(vb, ib)
    m.has_normals = nrm is not None
    m.has_uvs = uv is not None
    return m


def mesh_tensors(mesh: Mesh, stream=None) -> dict:
    """The mesh's vertex attributes as dense device tensors (positions/normals [V,3], uvs [V,2]
    float32) read out of its MeshVertex records by rt_mesh_unpack, plus `indices` [T,3] (int32 view
    of the index buffer, or None)."""
    lib = N.require_gpu()
    vb = mesh.vertices
    _check_device(vb, "Mesh.vertices")
    if vb.dtype != MeshVertex:
        raise TypeError("Mesh.vertices must be a buffer of MeshVertex")
    n = vb.shape[0]
    ts = stream if stream is not None else torch.cuda.current_stream()
    dev = torch.device("cuda")
    with torch.cuda.stream(ts):
        out = {"positions": torch.empty((n, 3), dtype=torch.float32, device=dev),
               "normals": torch.empty((n, 3), dtype=torch.float32, device=dev),
               "uvs": torch.empty((n, 2), dtype=torch.float32, device=dev)}
        N.check(lib.rt_mesh_unpack(vb.data_ptr(), n, out["positions"].data_ptr(),
                                   out["normals"].data_ptr(), out["uvs"].data_ptr(),
                                   N.stream_handle(ts)), "rt_mesh_unpack")
    ib = mesh.indices
    if ib is None:
        out["indices"] = None
    else:
        t = ib.tensor(torch.int64 if ib.dtype == np.int64 else torch.int32)
        out["indices"] = t.to(torch.int32).reshape(-1, 3)
    return out


class TextureBinding:
    """Textures and the mesh -> texture map of a scene uploaded once (device table of rt_texture
    and int32 map), for callers that render the same scene many times; pass it as `textures`."""

    def __init__(self, textures, mesh_textures):
        dev = torch.device("cuda")
        arr = (N.rt_texture * max(1, len(textures)))()
        for i, tx in enumerate(textures):
            _check_device(tx, "texture")
            if tx.dtype != torch.uint8 or tx.dim() != 3 or tx.shape[2] != 4 or not tx.is_contiguous():
                raise TypeError("texture tensors must be contiguous uint8 [H, W, 4]")
            arr[i].texels = tx.data_ptr()
            arr[i].width, arr[i].height = tx.shape[1], tx.shape[0]
        self.textures = list(textures)
        self.n_textures = len(textures)
        self.table = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8).to(dev)
        self.mesh_map = torch.tensor(list(mesh_textures), dtype=torch.int32).to(dev)


class Raycaster:
    """Closest-hit ray caster over a list of meshes (rendering/_raycaster.py:25-36)."""

    def __init__(self, models: typing.List[Mesh], stream=None):
        self.models = models
        self._stream = stream
        self._render_ws = None
        self._build_ads()

    # ------------------------------------------------------------------ build
    def _build_ads(self):
        lib = N.require_gpu()
        descs = (N.rt_mesh_desc * max(1, len(self.models)))()
        keep = []
        n_tris = 0
        for k, m in enumerate(self.models):
            _fill_mesh_desc(descs[k], m, keep)
            n_tris += descs[k].n_tris
        self.n_tris = n_tris
        self.mesh_tri_counts = [int(descs[k].n_tris) for k in range(len(self.models))]
        sizes = N.rt_scene_sizes()
        N.check(lib.rt_scene_sizes_for(n_tris, ctypes.byref(sizes)), "rt_scene_sizes_for")
        dev = torch.device("cuda")
        # allocate on the stream the build runs on: torch's allocator orders reuse of the freed
        # workspace after the build kernels on that stream
        # the scene's arrays are 256-byte-aligned slices of one allocation (one allocator call)
        parts = [("_nodes", sizes.nodes_bytes, torch.uint8), ("_nodes4", sizes.nodes4_bytes, torch.uint8),
                 ("_tris", sizes.tris_bytes, torch.uint8), ("_shade", sizes.shade_bytes, torch.uint8),
                 ("_bounds", 32, torch.float32), ("_mesh_offsets", 4 * (len(self.models) + 1), torch.int32),
                 ("keys", 4 * max(1, n_tris), torch.int32), ("order", 4 * max(1, n_tris), torch.int32)]
        offsets, total = [], 0
        for _, nbytes, _ in parts:
            offsets.append(total)
            total += (int(nbytes) + 255) // 256 * 256
        with torch.cuda.stream(self._torch_stream()):
            arena = torch.empty(total, dtype=torch.uint8, device=dev)
            ws = torch.empty(sizes.workspace_bytes, dtype=torch.uint8, device=dev)
        for (name, nbytes, dtype), off in zip(parts, offsets):
            setattr(self, name, arena[off:off + int(nbytes)].view(dtype))
        sc = N.rt_scene()
        sc.n_tris = n_tris
        sc.n_meshes = len(self.models)
        sc.nodes = self._nodes.data_ptr()
        sc.nodes4 = self._nodes4.data_ptr()
        sc.tris = self._tris.data_ptr()
        sc.shade = self._shade.data_ptr()
        sc.mesh_offsets = self._mesh_offsets.data_ptr()
        sc.bounds = self._bounds.data_ptr()
        code = lib.rt_scene_build(descs, len(self.models), ctypes.byref(sc), ws.data_ptr(),
                                  sizes.workspace_bytes, self.keys.data_ptr(),
                                  self.order.data_ptr(), N.stream_handle(self._stream))
        N.check(code, "rt_scene_build")
        self._scene = sc
        self._keep = keep
        del ws

    def device_bytes(self) -> int:
        """Device memory held by the built scene (BVH, triangle and shading records, keys)."""
        ts = [self._nodes, self._nodes4, self._tris, self._shade, self._bounds, self._mesh_offsets,
              self.keys, self.order]
        return int(sum(t.numel() * t.element_size() for t in ts))

    def release_inputs(self):
        """Drops the references to the input meshes' device buffers; queries and renders only read
        the built scene, so a resident scene need not keep its vertices."""
        self.n_models = len(self.models)
        self.models = [None] * self.n_models
        self._keep = []

    def _torch_stream(self):
        return self._stream if self._stream is not None else torch.cuda.current_stream()

    @property
    def root(self) -> int:
        return int(self._scene.root)

    def scene_arrays(self):
        """Raw device arrays of the built scene: nodes [n-1,16] f32, tris [n,12] f32,
        shade [n,16] f32, bounds [8], sorted Morton keys and primitive order."""
        n = self.n_tris
        ni = max(n - 1, 1)
        return {
            "nodes": self._nodes.view(torch.float32).view(ni, 16)[: max(n - 1, 0)],
            "nodes4": self._nodes4.view(torch.float32).view(ni, 32)[: max(n - 1, 0)],
            "tris": self._tris.view(torch.float32).view(max(n, 1), 12)[:n],
            "shade": self._shade.view(torch.float32).view(max(n, 1), 16)[:n],
            "bounds": self._bounds,
            "keys": self.keys[:n],
            "order": self.order[:n],
            "mesh_offsets": self._mesh_offsets,
        }

    # ------------------------------------------------------------------ queries
    def _rays_desc(self, rays, tmin, tmax):
        d = N.rt_rays()
        if isinstance(rays, DeviceArray):
            if rays.dtype != Ray:
                raise TypeError("rays must be a buffer of Ray")
            _check_device(rays, "rays")
            base = rays.data_ptr()
            n = rays.size
            d.ox, d.oy, d.oz = base, base + 4, base + 8
            d.dx, d.dy, d.dz = base + 16, base + 20, base + 24
            d.stride = Ray.itemsize
        elif isinstance(rays, dict):
            for k in ("ox", "oy", "oz", "dx", "dy", "dz"):
                _check_device(rays[k], "rays")
                if rays[k].dtype != torch.float32 or not rays[k].is_contiguous():
                    raise TypeError("SoA ray fields must be contiguous float32 tensors")
            n = rays["ox"].numel()
            d.ox, d.oy, d.oz = (rays[k].data_ptr() for k in ("ox", "oy", "oz"))
            d.dx, d.dy, d.dz = (rays[k].data_ptr() for k in ("dx", "dy", "dz"))
            d.stride = 4
            if "tmin" in rays:
                d.tmin = rays["tmin"].data_ptr()
            if "tmax" in rays:
                d.tmax = rays["tmax"].data_ptr()
        else:
            raise TypeError("rays must be a Ray buffer or a dict of SoA float32 tensors")
        d.tmin_default = float(tmin)
        d.tmax_default = float(tmax)
        return d, n

    def ray_cast(self, rays, tmin: float = 0.0, tmax: float = math.inf) -> DeviceArray:
        """Closest hit of every ray: a buffer of Hit (rendering/_raycaster.py:35-36)."""
        lib = N.require_gpu()
        d, n = self._rays_desc(rays, tmin, tmax)
        out = create_buffer(n, Hit)
        base = out.data_ptr()
        h = N.rt_hits()
        h.t, h.u, h.v = base, base + 4, base + 8
        h.index, h.mesh = base + 12, base + 16
        h.prim = None
        h.stride = Hit.itemsize
        N.check(lib.rt_trace_closest(ctypes.byref(self._scene), ctypes.byref(d), n, ctypes.byref(h),
                                     N.stream_handle(self._stream)), "rt_trace_closest")
        return out

    def trace_soa(self, rays: dict, tmin=0.0, tmax=math.inf, out=None):
        """Closest hit of SoA rays into SoA torch tensors t, u, v, prim, index, mesh."""
        lib = N.require_gpu()
        d, n = self._rays_desc(rays, tmin, tmax)
        dev = torch.device("cuda")
        if out is None:
            out = {k: torch.empty(n, dtype=torch.float32, device=dev) for k in ("t", "u", "v")}
            out.update({k: torch.empty(n, dtype=torch.int32, device=dev)
                        for k in ("prim", "index", "mesh")})
        h = N.rt_hits()
        for k in ("t", "u", "v", "prim", "index", "mesh"):
            setattr(h, k, out[k].data_ptr() if out.get(k) is not None else None)
        h.stride = 4
        N.check(lib.rt_trace_closest(ctypes.byref(self._scene), ctypes.byref(d), n, ctypes.byref(h),
                                     N.stream_handle(self._stream)), "rt_trace_closest")
        return out

    def occluded(self, rays, tmin: float = 0.0, tmax: float = math.inf) -> torch.Tensor:
        """Any-hit: uint8 tensor, 1 where some triangle lies in (tmin, tmax) along the ray."""
        lib = N.require_gpu()
        d, n = self._rays_desc(rays, tmin, tmax)
        out = torch.empty(max(n, 1), dtype=torch.uint8, device="cuda")
        N.check(lib.rt_trace_any(ctypes.byref(self._scene), ctypes.byref(d), n, out.data_ptr(),
                                 N.stream_handle(self._stream)), "rt_trace_any")
        return out[:n]

    @property
    def trace(self):
        """`rc.trace[n](rays, hits)` launch form of the reference's Dispatcher (rendering/_core.py:261)."""
        def _fn(n, rays, hits_out):
            res = self.trace_soa(rays, out=hits_out)
            return res
        return Dispatcher(_fn, "trace")

    # ------------------------------------------------------------------ camera rays / render
    @staticmethod
    def generate_rays(camera: Camera, width: int, height: int, stream=None) -> dict:
        """Primary rays of a width x height image as SoA float32 tensors (row-major pixels)."""
        lib = N.require_gpu()
        n = width * height
        dev = torch.device("cuda")
        out = {k: torch.empty(max(n, 1), dtype=torch.float32, device=dev)
               for k in ("ox", "oy", "oz", "dx", "dy", "dz")}
        cam = camera.to_c()
        N.check(lib.rt_generate_rays(ctypes.byref(cam), width, height,
                                     *[out[k].data_ptr() for k in ("ox", "oy", "oz", "dx", "dy", "dz")],
                                     N.stream_handle(stream)), "rt_generate_rays")
        return {k: v[:n] for k, v in out.items()}

    def cast_camera(self, camera: Camera, width: int, height: int, out=None, fields=("t", "u", "v", "prim")):
        """Closest hit of the primary rays of a width x height image (ray generation fused):
        SoA tensors over pixels (row-major), misses t=+inf, prim/index/mesh=-1."""
        lib = N.require_gpu()
        n = width * height
        dev = torch.device("cuda")
        if out is None:
            out = {}
            for k in fields:
                out[k] = torch.empty(max(n, 1), dtype=torch.float32 if k in ("t", "u", "v") else torch.int32,
                                     device=dev)
        h = N.rt_hits()
        for k in ("t", "u", "v", "prim", "index", "mesh"):
            setattr(h, k, out[k].data_ptr() if out.get(k) is not None else None)
        h.stride = 4
        cam = camera.to_c()
        N.check(lib.rt_trace_camera(ctypes.byref(self._scene), ctypes.byref(cam), width, height,
                                    ctypes.byref(h), N.stream_handle(self._stream)), "rt_trace_camera")
        return out

    def trace_paths(self, camera: Camera, width: int, height: int, spp: int = 4, bounces: int = 2,
                    seed: int = 0x5EED, frame: int = 0, albedo: float = 0.7, sky: float = 1.0,
                    ray_tmin: float = 1e-4, out=None, counts=None):
        """Diffuse path tracing under a uniform sky (rt_trace_paths): `spp` jittered camera rays
        per pixel, each followed by up to `bounces` cosine-weighted bounces (Philox4x32-10 with key
        (seed, frame)); returns the radiance image [height, width] float32. `counts` (optional
        int32 tensor of bounces+1) receives the rays traced at each depth."""
        lib = N.require_gpu()
        dev = torch.device("cuda")
        p = N.rt_paths_params(width, height, spp, bounces, seed & 0xFFFFFFFF, frame & 0xFFFFFFFF,
                              float(albedo), float(sky), float(ray_tmin))
        if out is None:
            out = torch.empty((height, width), dtype=torch.float32, device=dev)
        elif out.dtype != torch.float32 or out.numel() < width * height or not out.is_contiguous():
            raise TypeError("out must be a contiguous float32 tensor of height*width elements")
        need = lib.rt_paths_workspace_bytes(width, height, spp)
        if getattr(self, "_paths_ws", None) is None or self._paths_ws.numel() < need:
            self._paths_ws = torch.empty(need, dtype=torch.uint8, device=dev)
        o = N.rt_paths_out(out.data_ptr(), counts.data_ptr() if counts is not None else None)
        cam = camera.to_c()
        N.check(lib.rt_trace_paths(ctypes.byref(self._scene), ctypes.byref(cam), ctypes.byref(p),
                                   ctypes.byref(o), self._paths_ws.data_ptr(), self._paths_ws.numel(),
                                   N.stream_handle(self._stream)), "rt_trace_paths")
        return out

    def _render_workspace(self, n_pixels):
        lib = N.load()
        need = lib.rt_render_workspace_bytes(n_pixels)
        if self._render_ws is None or self._render_ws.numel() < need:
            self._render_ws = torch.empty(need, dtype=torch.uint8, device="cuda")
        return self._render_ws

    def _render_args(self, width, height, light, textures, mesh_textures, shadows, bilinear,
                     ambient, background, shadow_tmin, out, return_hits, stats):
        """rt_render_params / rt_render_out of one render plus the device tensors they point to."""
        n = width * height
        dev = torch.device("cuda")
        if light is None:
            light = np.array([1.0, 1.0, 1.0], dtype=np.float32) / np.float32(np.sqrt(3.0))
        light = np.asarray(light, dtype=np.float32)
        p = N.rt_render_params()
        p.width, p.height = width, height
        p.flags = (N.RT_RENDER_SHADOWS if shadows else 0) | (N.RT_RENDER_BILINEAR if bilinear else 0)
        p.light[:] = [float(v) for v in light]
        p.ambient = float(ambient)
        p.background[:] = [float(v) for v in background]
        p.shadow_tmin = float(shadow_tmin)
        tex_keep = []
        if isinstance(textures, TextureBinding):
            p.textures_dev = textures.table.data_ptr()
            p.n_textures = textures.n_textures
            p.mesh_texture_dev = textures.mesh_map.data_ptr()
            tex_keep.append(textures)
            textures, mesh_textures = None, None
        if textures:
            arr = (N.rt_texture * len(textures))()
            for i, tx in enumerate(textures):
                if isinstance(tx, tuple):
                    mem, desc = tx
                    if mem.dtype != RGBA:
                        raise TypeError("render samples RGBA (uchar4) textures")
                    arr[i].texels = texture_pointer(desc)
                    arr[i].width, arr[i].height = mem.shape[1], mem.shape[0]
                else:
                    _check_device(tx, "texture")
                    if tx.dtype != torch.uint8 or tx.dim() != 3 or tx.shape[2] != 4 or not tx.is_contiguous():
                        raise TypeError("texture tensors must be contiguous uint8 [H, W, 4]")
                    arr[i].texels = tx.data_ptr()
                    arr[i].width, arr[i].height = tx.shape[1], tx.shape[0]
            tdev = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8).to(dev)
            tex_keep.append(tdev)
            p.textures_dev = tdev.data_ptr()
            p.n_textures = len(textures)
            if mesh_textures is None:
                mesh_textures = [0] * len(self.models)
        if mesh_textures is not None:
            mt = torch.tensor(list(mesh_textures), dtype=torch.int32).to(dev)
            tex_keep.append(mt)
            p.mesh_texture_dev = mt.data_ptr()
        if out is None:
            out = torch.empty((height, width, 3), dtype=torch.float32, device=dev)
        elif out.dtype != torch.float32 or out.numel() < 3 * n or not out.is_contiguous():
            raise TypeError("out must be a contiguous float32 tensor of height*width*3 elements")
        o = N.rt_render_out()
        o.rgb = out.data_ptr()
        hits = None
        if return_hits:
            hits = {"t": torch.empty(max(n, 1), dtype=torch.float32, device=dev),
                    "prim": torch.empty(max(n, 1), dtype=torch.int32, device=dev)}
            o.t, o.prim = hits["t"].data_ptr(), hits["prim"].data_ptr()
        if stats is not None:
            o.stats = stats.data_ptr()
        return p, o, out, hits, tex_keep

    def render_many(self, requests, config=None):
        """Renders several images of this scene in one persistent launch (rt_render_batch):
        `requests` = list of dicts of render() keyword arguments (camera, width, height, ...).
        Returns the list of RGB images; see rendertoy_amd.batch for batches over many scenes."""
        from .batch import RenderRequest, BatchRender
        reqs = [RenderRequest(self, **r) for r in requests]
        b = BatchRender(reqs, config=config)
        b.launch()
        return b.images

    def render(self, camera: Camera, width: int, height: int, light=None, textures=None,
               mesh_textures=None, shadows=True, bilinear=True, ambient=0.2,
               background=(0.0, 0.0, 0.0), shadow_tmin=1e-4, out=None, return_hits=False,
               stats=None):
        """RGB float32 image [height, width, 3] of the scene: closest hit, Lambert with
        L = normalize(1,1,1) and ambient 0.2 by default (tutorials/lesson09_texture_mapping.py:73),
        shadow ray toward L, albedo from RGBA8 textures (texels as uint8 tensors [H, W, 4] or
        (memory, Texture2D) pairs from create_texture2D(w, h, RGBA)); `mesh_textures[m]` picks
        the texture of mesh m (-1: albedo 1)."""
        lib = N.require_gpu()
        n = width * height
        p, o, out, hits, tex_keep = self._render_args(width, height, light, textures, mesh_textures,
                                                      shadows, bilinear, ambient, background,
                                                      shadow_tmin, out, return_hits, stats)
        ws = self._render_workspace(n)
        cam = camera.to_c()
        N.check(lib.rt_render(ctypes.byref(self._scene), ctypes.byref(cam), ctypes.byref(p),
                              ctypes.byref(o), ws.data_ptr(), ws.numel(),
                              N.stream_handle(self._stream)), "rt_render")
        self._tex_keep = tex_keep  # alive until the next render on this stream
        if return_hits:
            return out, hits
        return out
