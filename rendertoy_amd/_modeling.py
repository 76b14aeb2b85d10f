"""Mesh input types of the ray-casting path (rendering/_modeling.py:8-34)."""
from enum import IntEnum

import numpy as np

from ._core import kernel_struct, float3, float2, create_buffer_from, DeviceArray


class WeldMode(IntEnum):
    NONE = 0
    POSITION = 1
    POSITION_NORMAL_TEXTURE = 2
    ALL_ATTRIBUTES = 3


class SubdivisionMode(IntEnum):
    NONE = 0
    FLAT = 1
    LOOP = 2
    BUTTERFLY = 3


@kernel_struct
class MeshVertex:
    """80-byte vertex record: P@0, N@16, C@32 (texture coordinate), T@48, B@64
    (rendering/_modeling.py:22-28)."""
    P: float3
    N: float3
    C: float2
    T: float3
    B: float3


class Mesh:
    """A triangle mesh: a device buffer of MeshVertex plus an optional index buffer (three vertex
    indices per triangle; None = consecutive vertex triples) (rendering/_modeling.py:31-34)."""

    def __init__(self, vertices: DeviceArray, indices: DeviceArray = None):
        self.vertices = vertices
        self.indices = indices

    @property
    def triangle_count(self) -> int:
        """rendering/_raycaster.py:33: V//3 without indices, I//3 with them."""
        if self.indices is None:
            return self.vertices.shape[0] // 3
        return self.indices.size // 3

    def clone(self) -> 'Mesh':
        raise Exception('Not implemented yet')

    def weld(self, weldMode: WeldMode, epsilon: float) -> 'Mesh':
        raise Exception('Not implemented yet')

    def simplify(self, max_vertices) -> 'Mesh':
        raise Exception('Not implemented yet')

    def subdivide(self, subdivision_mode: SubdivisionMode, **kwargs) -> 'Mesh':
        raise Exception('Not implemented yet')

    def compute_normals(self):
        """Area-weighted per-vertex normals, computed on the device and written into the N field
        of every vertex record (rendering/_modeling.py:48-49 is a stub in the reference); see
        rendertoy_amd._raycaster.compute_normals for the definition."""
        from ._raycaster import compute_normals
        compute_normals(self)

    def compute_tangents(self):
        raise Exception('Not implemented yet')


def mesh_from_arrays(positions, indices=None, normals=None, uvs=None) -> Mesh:
    """Packs host arrays (positions [V,3], optional normals [V,3], uvs [V,2], indices [T,3]) into
    MeshVertex/int32 device buffers."""
    positions = np.asarray(positions, dtype=np.float32).reshape(-1, 3)
    v = np.zeros(positions.shape[0], dtype=MeshVertex)
    flat = v.view(np.float32).reshape(positions.shape[0], MeshVertex.itemsize // 4)
    flat[:, 0:3] = positions
    if normals is not None:
        flat[:, 4:7] = np.asarray(normals, dtype=np.float32).reshape(-1, 3)
    if uvs is not None:
        flat[:, 8:10] = np.asarray(uvs, dtype=np.float32).reshape(-1, 2)
    vb = create_buffer_from(v)
    ib = None
    if indices is not None:
        ib = create_buffer_from(np.asarray(indices, dtype=np.int32).reshape(-1))
    m = Mesh(vb, ib)
    m.has_normals = normals is not None
    m.has_uvs = uvs is not None
    return m
