"""rendertoy_amd — MI355X-native drop-in for the raycaster path of the rendertoy `rendering`
package (`import rendertoy_amd as ren` in place of `import rendering as ren`).

Exports the names of rendering/__init__.py:1-15 that the ray-casting path uses (buffers, structs,
vector types, host math, Mesh/MeshVertex, Texture2D) plus the ray caster itself, which the
reference defines in rendering/_raycaster.py but does not export.
"""
from ._core import (kernel_struct, create_buffer_from, create_buffer, create_struct, mapped,
                    create_struct_from, DeviceArray, Buffer, make_float2, make_float3, make_float4,
                    make_float4x4, make_int2, make_int3, make_int4, translate, identity, scale,
                    rotate, matmul, to_array, clear, perspective, look_at, normalize, dot, cross,
                    direction, Dispatcher, struct_layout)
from ._core import (float2, float3, float4, int2, int3, int4, uint2, uint3, uint4, float4x4, RGBA,
                    Texture2D, create_texture2D, texture_pointer)
from ._modeling import Mesh, WeldMode, SubdivisionMode, MeshVertex, mesh_from_arrays
from ._raycaster import Raycaster, BVH_AABB, BVH_Triangle, Ray, Hit, Camera, compute_normals, \
    mesh_from_tensors, mesh_tensors
from ._loaders import load_obj
from .tuning import get_tuning, set_tuning, tuning

__all__ = [n for n in dir() if not n.startswith("_")]
