"""The ray caster of rendering/_raycaster.py on the MI355X kernels of librtx_hip.so.

Reference API kept: `Raycaster(models: List[Mesh])` builds the acceleration structure
(`_build_ads`, rendering/_raycaster.py:26-33) and `ray_cast(rays) -> array` answers closest-hit
queries (:35-36); the `BVH_AABB` and `BVH_Triangle` structs (:8-22) are registered with the same
layout, and `export_ads()` fills them from the built hierarchy. Additions: `Ray`/`Hit` structs,
`occluded(rays)`, `render(...)` and `generate_rays(...)`. Triangle identity of a hit is
(mesh, index) as BVH_Triangle defines it (:16-22).

Streams: a Raycaster made with `stream=s` runs its build and every query and render on `s`, after
the work already queued on the caller's current stream (inputs made there are ready); its outputs
are ready on `s` (wait for `rc.stream` before reading them on another stream).
Every call runs on the GPU; there is no CPU path.
"""
import contextlib
import ctypes
import math
import typing

import numpy as np
import torch

from . import _native as N
from .tuning import make_tuning, as_dict as tuning_dict, tuning_arg
from ._core import (kernel_struct, float3, float4, float4x4, DeviceArray, create_buffer, Dispatcher,
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

    @classmethod
    def from_matrices(cls, view, proj):
        """The camera of a view matrix from look_at(eye, target, up) and a projection from
        perspective(fov, aspect, znear, zfar) (rendering/_core.py:528-548; the matrices the
        reference's shaders take, tutorials/lesson08_rasterization.py:94-99): the basis is read
        from the view matrix's columns (x, y, z axes), the eye from its translation row, and
        tan(fov/2), aspect from the projection's diagonal. The ray of pixel (col, row) is then the
        inverse of that projection through the pixel centre."""
        V = np.asarray(to_array(view) if np.asarray(view).dtype == float4x4 else view, np.float64)
        P = np.asarray(to_array(proj) if np.asarray(proj).dtype == float4x4 else proj, np.float64)
        V, P = V.reshape(4, 4), P.reshape(4, 4)
        if P[2, 3] != 1.0 or P[3, 3] != 0.0 or P[0, 0] <= 0 or P[1, 1] <= 0:
            raise ValueError("proj must be a perspective() matrix (w = view z, positive scales)")
        M = V[:3, :3]  # columns: x, y, z axes
        cam = cls.__new__(cls)
        cam.eye = (-(V[3, :3] @ M.T)).astype(np.float32)
        ws, hs = np.float32(P[0, 0]), np.float32(P[1, 1])
        cam.right = (V[:3, 0].astype(np.float32) / ws).astype(np.float32)
        cam.up = (V[:3, 1].astype(np.float32) / hs).astype(np.float32)
        cam.forward = V[:3, 2].astype(np.float32)
        cam.fov = float(2.0 * math.atan(1.0 / float(hs)))
        cam.aspect = float(hs / ws)
        return cam

    def as_array(self):
        """12 float32: eye, right, up, forward."""
        return np.concatenate([self.eye, self.right, self.up, self.forward]).astype(np.float32)

    def to_c(self):
        c = N.rt_camera()
        for name in ("eye", "right", "up", "forward"):
            getattr(c, name)[:] = [float(v) for v in getattr(self, name)]
        return c


# the default light direction, normalize(1, 1, 1) in float32 (tutorials/lesson09_texture_mapping.py:73)
_DEFAULT_LIGHT = [float(v) for v in np.array([1.0, 1.0, 1.0], dtype=np.float32) / np.float32(np.sqrt(3.0))]


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


def _check_out(t, what, dtype, n):
    """An output tensor the kernels write n elements of: device, dtype, contiguous, large enough."""
    _check_device(t, what)
    if t.dtype != dtype or not t.is_contiguous():
        raise TypeError(f"{what} must be a contiguous {dtype} device tensor")
    if t.numel() < n:
        raise ValueError(f"{what} holds {t.numel()} elements, {n} are written")


def _narrow_indices(ib, n_vertices):
    """int32 device copy of an index buffer of another integer type. Values outside
    [0, n_vertices) become n_vertices - 1, the rule the build applies to int32 indices, so no
    value can wrap into range."""
    kinds = {np.dtype(np.int64): torch.int64, np.dtype(np.int16): torch.int16,
             np.dtype(np.int8): torch.int8, np.dtype(np.uint8): torch.uint8}
    dt = np.dtype(ib.dtype)
    if dt in kinds:
        t = ib.tensor(kinds[dt]).to(torch.int64)
    elif dt == np.uint16:
        t = ib.tensor(torch.int16).to(torch.int64) & 0xFFFF
    else:
        raise TypeError(f"index buffers must be int32, uint32, int64, int16, uint16, int8 or uint8 "
                        f"(got {dt})")
    last = max(int(n_vertices) - 1, 0)
    return torch.where((t < 0) | (t >= n_vertices), torch.full_like(t, last), t).to(torch.int32)


def _fill_mesh_desc(d, m, keep):
    """Fills rt_mesh_desc `d` from Mesh `m` (MeshVertex records, optional index buffer); narrowed
    index copies are appended to `keep`, which must outlive the call that reads them (they are made
    on the current stream)."""
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
        if np.dtype(ib.dtype) in (np.dtype(np.int32), np.dtype(np.uint32)):
            # uint32 indices are read as their bits: values >= 2^31 are out of range either way
            d.indices = ib.data_ptr()
        else:
            # e.g. the int64 index buffers of load_obj (rendering/_loaders.py:17)
            ib32 = _narrow_indices(ib, d.n_vertices)
            keep.append(ib32)
            d.indices = ib32.data_ptr()
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


def _texture_rgba8(tx, keep, stream):
    """(device pointer, width, height) of RGBA8 texels for one texture argument: a uint8 tensor
    [H, W, 4], a float32 tensor [H, W, 4], or a (memory, Texture2D) pair from create_texture2D with
    RGBA or float4 texels (the reference's default, rendering/_core.py:557-578). Float texels are
    converted once, here, with the UNORM_INT8 rule (rt_texels_to_rgba8). Whatever device memory
    the returned pointer (or the conversion) reads is appended to `keep`: a caller that relaunches
    later (a batch, a captured graph) holds it alive."""
    if isinstance(tx, tuple):
        mem, desc = tx
        h, w = mem.shape[0], mem.shape[1]
        keep.append(tx)
        if mem.dtype == RGBA:
            return texture_pointer(desc), w, h
        if mem.dtype != float4:
            raise TypeError("texture memory must hold RGBA (uchar4) or float4 texels")
        src, n = texture_pointer(desc), w * h
    else:
        _check_device(tx, "texture")
        if tx.dim() != 3 or tx.shape[2] != 4 or not tx.is_contiguous():
            raise TypeError("texture tensors must be contiguous [H, W, 4]")
        h, w = tx.shape[0], tx.shape[1]
        if tx.dtype not in (torch.uint8, torch.float32):
            raise TypeError("texture tensors must be uint8 (RGBA8) or float32 (RGBA) [H, W, 4]")
        keep.append(tx)
        if tx.dtype == torch.uint8:
            return tx.data_ptr(), w, h
        src, n = tx.data_ptr(), w * h
    lib = N.load()
    dst = torch.empty((h, w, 4), dtype=torch.uint8, device=torch.device("cuda"))
    N.check(lib.rt_texels_to_rgba8(src, n, dst.data_ptr(), N.stream_handle(stream)), "rt_texels_to_rgba8")
    keep.append(dst)
    return dst.data_ptr(), w, h


class TextureBinding:
    """Textures and the mesh -> texture map of a scene uploaded once (device table of rt_texture
    and int32 map), for callers that render the same scene many times; pass it as `textures`.
    Float texels are converted to RGBA8 once, here. `mesh_textures[m]` is the texture of mesh m
    (-1: albedo 1); it must have one entry per mesh of the scenes it is used with."""

    def __init__(self, textures, mesh_textures, stream=None):
        dev = torch.device("cuda")
        ts = stream if stream is not None else torch.cuda.current_stream()
        arr = (N.rt_texture * max(1, len(textures)))()
        self._keep = []
        with torch.cuda.stream(ts):
            for i, tx in enumerate(textures):
                arr[i].texels, arr[i].width, arr[i].height = _texture_rgba8(tx, self._keep, ts)
            self.table = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8).to(dev)
            self.mesh_map = torch.tensor(list(mesh_textures), dtype=torch.int32).to(dev)
        self.textures = list(textures)
        self.n_textures = len(textures)
        self.n_meshes = len(list(mesh_textures))


class Raycaster:
    """Closest-hit ray caster over a list of meshes (rendering/_raycaster.py:25-36).

    `stream`: the torch stream its kernels run on (default: the caller's current stream at each
    call). `tuning`: launch choices for this scene (a dict of rendertoy_amd.tuning fields over the
    process default at construction, e.g. {"builder": "ploc", "query": "binary"}); every method also
    takes a `tuning=` dict for one call. Without either, each call uses the process default."""

    def __init__(self, models: typing.List[Mesh], stream=None, tuning=None):
        self.models = models
        self._stream = stream
        self._tuning = None if tuning is None else make_tuning(None, **tuning)
        self._build_ads()

    # ------------------------------------------------------------------ helpers
    @property
    def stream(self):
        """The stream the kernels of this scene run on."""
        return self._torch_stream()

    def _torch_stream(self):
        return self._stream if self._stream is not None else torch.cuda.current_stream()

    @contextlib.contextmanager
    def _on_stream(self):
        """Runs the block with this scene's stream current (allocations, copies and kernels all on
        it), after the work already queued on the caller's stream."""
        ts = self._torch_stream()
        cur = torch.cuda.current_stream()
        if ts != cur:
            ts.wait_stream(cur)
        with torch.cuda.stream(ts):
            yield ts

    def _tu(self, tuning):
        """The rt_tuning of one call (None: this scene's, or the process default)."""
        if tuning is None:
            return self._tuning
        base = None if self._tuning is None else tuning_dict(self._tuning)
        return make_tuning(base, **tuning)

    # ------------------------------------------------------------------ build
    def _build_ads(self):
        lib = N.require_gpu()
        descs = (N.rt_mesh_desc * max(1, len(self.models)))()
        keep = []
        n_tris = 0
        with self._on_stream() as ts:
            for k, m in enumerate(self.models):
                _fill_mesh_desc(descs[k], m, keep)
                n_tris += descs[k].n_tris
            self.n_tris = n_tris
            self.mesh_tri_counts = [int(descs[k].n_tris) for k in range(len(self.models))]
            sizes = N.rt_scene_sizes()
            N.check(lib.rt_scene_sizes_for(n_tris, ctypes.byref(sizes)), "rt_scene_sizes_for")
            dev = torch.device("cuda")
            # the scene's arrays are 256-byte-aligned slices of one allocation (one allocator
            # call); allocated on the build's stream, so torch orders reuse of the freed workspace
            # after the build kernels
            parts = [("_nodes", sizes.nodes_bytes, torch.uint8), ("_nodes4", sizes.nodes4_bytes, torch.uint8),
                     ("_tris", sizes.tris_bytes, torch.uint8), ("_shade", sizes.shade_bytes, torch.uint8),
                     ("_bounds", 32, torch.float32), ("_mesh_offsets", 4 * (len(self.models) + 1), torch.int32),
                     ("_prim_slot", sizes.prim_slot_bytes, torch.int32),
                     ("keys", 4 * max(1, n_tris), torch.int32), ("order", 4 * max(1, n_tris), torch.int32)]
            offsets, total = [], 0
            for _, nbytes, _ in parts:
                offsets.append(total)
                total += (int(nbytes) + 255) // 256 * 256
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
            sc.prim_slot = self._prim_slot.data_ptr()
            code = lib.rt_scene_build(descs, len(self.models), ctypes.byref(sc), ws.data_ptr(),
                                      sizes.workspace_bytes, self.keys.data_ptr(),
                                      self.order.data_ptr(), tuning_arg(self._tuning),
                                      N.stream_handle(ts))
            N.check(code, "rt_scene_build")
            del ws
        self._scene = sc
        self._keep = keep

    def device_bytes(self) -> int:
        """Device memory held by the built scene (BVH, triangle and shading records, keys)."""
        ts = [self._nodes, self._nodes4, self._tris, self._shade, self._bounds, self._mesh_offsets,
              self._prim_slot, self.keys, self.order]
        return int(sum(t.numel() * t.element_size() for t in ts))

    def release_inputs(self):
        """Drops the references to the input meshes' device buffers; queries and renders only read
        the built scene, so a resident scene need not keep its vertices (export_ads needs them)."""
        self.n_models = len(self.models)
        self.models = [None] * self.n_models
        self._keep = []

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
            "prim_slot": self._prim_slot[:n],
            "mesh_offsets": self._mesh_offsets,
        }

    def export_ads(self):
        """The built hierarchy in the reference's record types (rendering/_raycaster.py:8-22):
        (nodes, triangles) as device buffers of BVH_AABB and BVH_Triangle. nodes[0] is the root
        (box = scene bounds); an internal node has count 0 and its children at `elements` and
        `elements` + 1; a leaf has count 1..4 triangles at triangles[elements : elements + count].
        Triangles are in leaf order with the hit identity (index within the mesh, mesh position),
        and their vertices are the input meshes' (so the meshes must not have been released).
        Boxes are the hierarchy's conservative child boxes: every triangle lies inside each box on
        its path. Synchronises once (the node count)."""
        lib = N.require_gpu()
        if any(m is None for m in self.models):
            raise RuntimeError("export_ads needs the input meshes; release_inputs() dropped them")
        descs = (N.rt_mesh_desc * max(1, len(self.models)))()
        keep = []
        n = self.n_tris
        with self._on_stream() as ts:
            for k, m in enumerate(self.models):
                _fill_mesh_desc(descs[k], m, keep)
            nodes = create_buffer(max(2 * n - 1, 1), BVH_AABB)
            tris = create_buffer(max(n, 1), BVH_Triangle)
            count = torch.zeros(1, dtype=torch.int32, device=torch.device("cuda"))
            N.check(lib.rt_scene_export(ctypes.byref(self._scene), descs, len(self.models),
                                        nodes.data_ptr(), tris.data_ptr(), count.data_ptr(),
                                        N.stream_handle(ts)), "rt_scene_export")
            c = int(count.item())
        return nodes[:c], tris[:n]

    # ------------------------------------------------------------------ queries
    def _rays_desc(self, rays, tmin, tmax, n=None):
        """rt_rays of a Ray buffer or a dict of SoA float32 tensors (ox, oy, oz, dx, dy, dz and
        optional per-ray tmin, tmax); `n`: trace only the first n rays (default: all)."""
        d = N.rt_rays()
        if isinstance(rays, DeviceArray):
            if rays.dtype != Ray:
                raise TypeError("rays must be a buffer of Ray")
            _check_device(rays, "rays")
            base = rays.data_ptr()
            avail = rays.size
            d.ox, d.oy, d.oz = base, base + 4, base + 8
            d.dx, d.dy, d.dz = base + 16, base + 20, base + 24
            d.stride = Ray.itemsize
        elif isinstance(rays, dict):
            missing = [k for k in ("ox", "oy", "oz", "dx", "dy", "dz") if k not in rays]
            if missing:
                raise KeyError(f"SoA rays lack the fields {missing}")
            avail = rays["ox"].numel()
            for k in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax"):
                if k not in rays or rays[k] is None:
                    continue
                t = rays[k]
                _check_device(t, f"rays[{k!r}]")
                if t.dtype != torch.float32 or not t.is_contiguous():
                    raise TypeError("SoA ray fields must be contiguous float32 tensors")
                if t.numel() != avail:
                    raise ValueError(f"rays[{k!r}] has {t.numel()} elements, rays['ox'] {avail}")
            d.ox, d.oy, d.oz = (rays[k].data_ptr() for k in ("ox", "oy", "oz"))
            d.dx, d.dy, d.dz = (rays[k].data_ptr() for k in ("dx", "dy", "dz"))
            d.stride = 4
            if rays.get("tmin") is not None:
                d.tmin = rays["tmin"].data_ptr()
            if rays.get("tmax") is not None:
                d.tmax = rays["tmax"].data_ptr()
        else:
            raise TypeError("rays must be a Ray buffer or a dict of SoA float32 tensors")
        if n is None:
            n = avail
        if not 0 <= int(n) <= avail:
            raise ValueError(f"{n} rays requested, the buffer holds {avail}")
        if n >= 2 ** 32:
            raise ValueError("at most 2^32 - 1 rays per call")
        d.tmin_default = float(tmin)
        d.tmax_default = float(tmax)
        return d, int(n)

    def ray_cast(self, rays, tmin: float = 0.0, tmax: float = math.inf, tuning=None) -> DeviceArray:
        """Closest hit of every ray: a buffer of Hit (rendering/_raycaster.py:35-36)."""
        lib = N.require_gpu()
        d, n = self._rays_desc(rays, tmin, tmax)
        with self._on_stream() as ts:
            out = create_buffer(n, Hit)
            base = out.data_ptr()
            h = N.rt_hits()
            h.t, h.u, h.v = base, base + 4, base + 8
            h.index, h.mesh = base + 12, base + 16
            h.prim = None
            h.stride = Hit.itemsize
            N.check(lib.rt_trace_closest(ctypes.byref(self._scene), ctypes.byref(d), n, ctypes.byref(h),
                                         tuning_arg(self._tu(tuning)), N.stream_handle(ts)),
                    "rt_trace_closest")
        return out

    _HIT_FIELDS = (("t", torch.float32), ("u", torch.float32), ("v", torch.float32),
                   ("prim", torch.int32), ("index", torch.int32), ("mesh", torch.int32))

    def trace_soa(self, rays: dict, tmin=0.0, tmax=math.inf, out=None, n=None, tuning=None):
        """Closest hit of SoA rays into SoA torch tensors t, u, v, prim, index, mesh (`out`: a dict
        of any of them, each at least n elements; default: all six, allocated). `n`: trace only
        the first n rays."""
        lib = N.require_gpu()
        d, n = self._rays_desc(rays, tmin, tmax, n)
        with self._on_stream() as ts:
            if out is None:
                out = {k: torch.empty(n, dtype=dt, device=torch.device("cuda")) for k, dt in self._HIT_FIELDS}
            h = N.rt_hits()
            for k, dt in self._HIT_FIELDS:
                if out.get(k) is not None:
                    _check_out(out[k], f"out[{k!r}]", dt, n)
                setattr(h, k, out[k].data_ptr() if out.get(k) is not None else None)
            h.stride = 4
            N.check(lib.rt_trace_closest(ctypes.byref(self._scene), ctypes.byref(d), n, ctypes.byref(h),
                                         tuning_arg(self._tu(tuning)), N.stream_handle(ts)),
                    "rt_trace_closest")
        return out

    def occluded(self, rays, tmin: float = 0.0, tmax: float = math.inf, out=None, n=None,
                 tuning=None) -> torch.Tensor:
        """Any-hit: uint8 tensor, 1 where some triangle lies in (tmin, tmax) along the ray."""
        lib = N.require_gpu()
        d, n = self._rays_desc(rays, tmin, tmax, n)
        with self._on_stream() as ts:
            if out is None:
                out = torch.empty(max(n, 1), dtype=torch.uint8, device="cuda")
            else:
                _check_out(out, "out", torch.uint8, n)
            N.check(lib.rt_trace_any(ctypes.byref(self._scene), ctypes.byref(d), n, out.data_ptr(),
                                     tuning_arg(self._tu(tuning)), N.stream_handle(ts)),
                    "rt_trace_any")
        return out[:n]

    @property
    def trace(self):
        """`rc.trace[n](rays, hits)`: the launch form of the reference's Dispatcher
        (rendering/_core.py:252-253, 261-290): exactly the first n rays are traced (the reference
        launches n work-items and every id >= n returns); `hits` is a dict of SoA output tensors
        of at least n elements (None: allocated)."""
        def _fn(n, rays, hits_out=None):
            return self.trace_soa(rays, out=hits_out, n=n)
        return Dispatcher(_fn, "trace")

    # ------------------------------------------------------------------ camera rays / render
    @staticmethod
    def generate_rays(camera: Camera, width: int, height: int, stream=None) -> dict:
        """Primary rays of a width x height image as SoA float32 tensors (row-major pixels)."""
        lib = N.require_gpu()
        n = width * height
        dev = torch.device("cuda")
        ts = stream if stream is not None else torch.cuda.current_stream()
        with torch.cuda.stream(ts):
            out = {k: torch.empty(max(n, 1), dtype=torch.float32, device=dev)
                   for k in ("ox", "oy", "oz", "dx", "dy", "dz")}
            cam = camera.to_c()
            N.check(lib.rt_generate_rays(ctypes.byref(cam), width, height,
                                         *[out[k].data_ptr() for k in ("ox", "oy", "oz", "dx", "dy", "dz")],
                                         N.stream_handle(ts)), "rt_generate_rays")
        return {k: v[:n] for k, v in out.items()}

    def cast_camera(self, camera: Camera, width: int, height: int, out=None,
                    fields=("t", "u", "v", "prim"), tuning=None):
        """Closest hit of the primary rays of a width x height image (ray generation fused):
        SoA tensors over pixels (row-major), misses t=+inf, prim/index/mesh=-1. `out`: a dict of
        any of t, u, v, prim, index, mesh tensors of at least width*height elements."""
        lib = N.require_gpu()
        n = width * height
        with self._on_stream() as ts:
            if out is None:
                dts = dict(self._HIT_FIELDS)
                out = {k: torch.empty(max(n, 1), dtype=dts[k], device=torch.device("cuda")) for k in fields}
            h = N.rt_hits()
            for k, dt in self._HIT_FIELDS:
                if out.get(k) is not None:
                    _check_out(out[k], f"out[{k!r}]", dt, n)
                setattr(h, k, out[k].data_ptr() if out.get(k) is not None else None)
            h.stride = 4
            cam = camera.to_c()
            ws = self._camera_workspace(width, height)
            N.check(lib.rt_trace_camera(ctypes.byref(self._scene), ctypes.byref(cam), width, height,
                                        ctypes.byref(h), None if ws is None else ws.data_ptr(),
                                        0 if ws is None else ws.numel(), tuning_arg(self._tu(tuning)),
                                        N.stream_handle(ts)), "rt_trace_camera")
        return out

    def trace_paths(self, camera: Camera, width: int, height: int, spp: int = 4, bounces: int = 2,
                    seed: int = 0x5EED, frame: int = 0, albedo: float = 0.7, sky: float = 1.0,
                    ray_tmin: float = 1e-4, out=None, counts=None, tuning=None):
        """Diffuse path tracing under a uniform sky (rt_trace_paths): `spp` jittered camera rays
        per pixel, each followed by up to `bounces` cosine-weighted bounces (Philox4x32-10 with key
        (seed, frame)); returns the radiance image [height, width] float32. `counts` (optional
        int32 tensor of bounces+1) receives the rays traced at each depth."""
        lib = N.require_gpu()
        p = N.rt_paths_params(width, height, spp, bounces, seed & 0xFFFFFFFF, frame & 0xFFFFFFFF,
                              float(albedo), float(sky), float(ray_tmin))
        with self._on_stream() as ts:
            if out is None:
                out = torch.empty((height, width), dtype=torch.float32, device=torch.device("cuda"))
            else:
                _check_out(out, "out", torch.float32, width * height)
            if counts is not None:
                _check_out(counts, "counts", torch.int32, bounces + 1)
            need = lib.rt_paths_workspace_bytes(width, height, spp)
            if getattr(self, "_paths_ws", None) is None or self._paths_ws.numel() < need:
                self._paths_ws = torch.empty(need, dtype=torch.uint8, device=torch.device("cuda"))
            o = N.rt_paths_out(out.data_ptr(), counts.data_ptr() if counts is not None else None)
            cam = camera.to_c()
            N.check(lib.rt_trace_paths(ctypes.byref(self._scene), ctypes.byref(cam), ctypes.byref(p),
                                       ctypes.byref(o), self._paths_ws.data_ptr(), self._paths_ws.numel(),
                                       tuning_arg(self._tu(tuning)), N.stream_handle(ts)),
                    "rt_trace_paths")
        return out

    def _camera_workspace(self, width, height):
        """The camera kernel's workspace on the current stream, or None when it needs none. One
        per stream: calls on one stream run in order, calls on two streams must not share one."""
        need = N.load().rt_trace_camera_workspace_bytes(width, height)
        if need == 0:
            return None
        pool = self.__dict__.setdefault("_cam_ws", {})
        key = torch.cuda.current_stream().cuda_stream
        ws = pool.get(key)
        if ws is None or ws.numel() < need:
            ws = pool[key] = torch.empty(need, dtype=torch.uint8, device="cuda")
        return ws

    def _render_workspace(self, n_pixels):
        """rt_render's workspace on the current stream (the partial counts of a render that returns
        counts). One per stream: calls on one stream run in order, calls on two streams must not
        share one."""
        need = N.load().rt_render_workspace_bytes(n_pixels)
        pool = self.__dict__.setdefault("_render_ws_pool", {})
        key = torch.cuda.current_stream().cuda_stream
        ws = pool.get(key)
        if ws is None or ws.numel() < need:
            ws = pool[key] = torch.empty(need, dtype=torch.uint8, device="cuda")
        return ws

    @property
    def n_meshes(self):
        return len(self.models)

    def _render_args(self, width, height, light, textures, mesh_textures, shadows, bilinear,
                     ambient, background, shadow_tmin, out, return_hits, stats):
        """rt_render_params / rt_render_out of one render plus the device tensors they point to
        (allocated on the current stream)."""
        n = width * height
        dev = torch.device("cuda")
        p = N.rt_render_params()
        p.width, p.height = width, height
        p.flags = (N.RT_RENDER_SHADOWS if shadows else 0) | (N.RT_RENDER_BILINEAR if bilinear else 0)
        p.light[:] = _DEFAULT_LIGHT if light is None else [float(v) for v in np.asarray(light, dtype=np.float32)]
        p.ambient = float(ambient)
        p.background[:] = [float(v) for v in background]
        p.shadow_tmin = float(shadow_tmin)
        tex_keep = []
        if isinstance(textures, TextureBinding):
            if textures.n_meshes != self.n_meshes:
                raise ValueError(f"the texture binding maps {textures.n_meshes} meshes, the scene has "
                                 f"{self.n_meshes}")
            p.textures_dev = textures.table.data_ptr()
            p.n_textures = textures.n_textures
            p.mesh_texture_dev = textures.mesh_map.data_ptr()
            tex_keep.append(textures)
            textures, mesh_textures = None, None
        if textures:
            arr = (N.rt_texture * len(textures))()
            ts = torch.cuda.current_stream()
            for i, tx in enumerate(textures):
                arr[i].texels, arr[i].width, arr[i].height = _texture_rgba8(tx, tex_keep, ts)
            tdev = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8).to(dev)
            tex_keep.append(tdev)
            p.textures_dev = tdev.data_ptr()
            p.n_textures = len(textures)
            if mesh_textures is None:
                mesh_textures = [0] * self.n_meshes
        if mesh_textures is not None:
            mesh_textures = list(mesh_textures)
            if len(mesh_textures) != self.n_meshes:
                raise ValueError(f"mesh_textures has {len(mesh_textures)} entries, the scene has "
                                 f"{self.n_meshes} meshes")
            mt = torch.tensor(mesh_textures, dtype=torch.int32).to(dev)
            tex_keep.append(mt)
            p.mesh_texture_dev = mt.data_ptr()
        if out is None:
            out = torch.empty((height, width, 3), dtype=torch.float32, device=dev)
        else:
            _check_out(out, "out", torch.float32, 3 * n)
        o = N.rt_render_out()
        o.rgb = out.data_ptr()
        hits = None
        if return_hits:
            hits = {"t": torch.empty(max(n, 1), dtype=torch.float32, device=dev),
                    "prim": torch.empty(max(n, 1), dtype=torch.int32, device=dev)}
            o.t, o.prim = hits["t"].data_ptr(), hits["prim"].data_ptr()
        if stats is not None:
            _check_out(stats, "stats", torch.int32, 4)
            o.stats = stats.data_ptr()
        return p, o, out, hits, tex_keep

    def render_many(self, requests, config=None):
        """Renders several images of this scene in one persistent launch (rt_render_batch):
        `requests` = list of dicts of render() keyword arguments (camera, width, height, ...).
        Returns the list of RGB images; see rendertoy_amd.batch for batches over many scenes."""
        from .batch import RenderRequest, BatchRender
        reqs = [RenderRequest(self, **r) for r in requests]
        b = BatchRender(reqs, config=config, stream=self._stream)
        b.launch()
        return b.images

    def render(self, camera: Camera, width: int, height: int, light=None, textures=None,
               mesh_textures=None, shadows=True, bilinear=True, ambient=0.2,
               background=(0.0, 0.0, 0.0), shadow_tmin=1e-4, out=None, return_hits=False,
               stats=None, tuning=None):
        """RGB float32 image [height, width, 3] of the scene: closest hit, Lambert with
        L = normalize(1,1,1) and ambient 0.2 by default (tutorials/lesson09_texture_mapping.py:73),
        shadow ray toward L, albedo from the textures: uint8 RGBA8 or float32 RGBA tensors
        [H, W, 4], (memory, Texture2D) pairs from create_texture2D(w, h) (float4 texels, the
        reference's default, converted to RGBA8 with the UNORM rule) or create_texture2D(w, h,
        RGBA), or a TextureBinding; `mesh_textures[m]` picks the texture of mesh m (-1: albedo 1)."""
        lib = N.require_gpu()
        n = width * height
        with self._on_stream() as ts:
            p, o, out, hits, tex_keep = self._render_args(width, height, light, textures, mesh_textures,
                                                          shadows, bilinear, ambient, background,
                                                          shadow_tmin, out, return_hits, stats)
            ws = self._render_workspace(n)
            cam = camera.to_c()
            N.check(lib.rt_render(ctypes.byref(self._scene), ctypes.byref(cam), ctypes.byref(p),
                                  ctypes.byref(o), ws.data_ptr(), ws.numel(),
                                  tuning_arg(self._tu(tuning)), N.stream_handle(ts)), "rt_render")
        self._tex_keep = tex_keep  # alive until the next render on this stream
        if return_hits:
            return out, hits
        return out
