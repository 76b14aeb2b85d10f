"""Host-side API of rendertoy_amd (no GPU): the rendering/_core.py buffer, struct and math
conventions it mirrors."""
import math

import numpy as np
import pytest

import rendertoy_amd as ren


def test_vector_dtypes_follow_opencl_layout():
    assert ren.float2.itemsize == 8 and ren.float3.itemsize == 16 and ren.float4.itemsize == 16
    assert ren.float4x4.itemsize == 64 and ren.int3.itemsize == 16 and ren.RGBA.itemsize == 4
    v = ren.make_float3(1, 2, 3)
    assert v["x"] == 1 and v["y"] == 2 and v["z"] == 3
    np.testing.assert_array_equal(ren.to_array(v), [1, 2, 3])
    np.testing.assert_array_equal(ren.to_array(ren.make_float4(1, 2, 3, 4)), [1, 2, 3, 4])


def test_struct_layouts_of_the_path():
    # MeshVertex (rendering/_modeling.py:22-28; offsets evidenced by _modeling.py:75,78,84)
    assert ren.struct_layout(ren.MeshVertex) == [("P", 0, 16), ("N", 16, 16), ("C", 32, 8),
                                                 ("T", 48, 16), ("B", 64, 16)]
    assert ren.MeshVertex.itemsize == 80
    # BVH_AABB / BVH_Triangle (rendering/_raycaster.py:8-22) by OpenCL C alignment rules
    assert ren.struct_layout(ren.BVH_AABB) == [("min", 0, 16), ("max", 16, 16), ("count", 32, 4),
                                               ("elements", 36, 4)]
    assert ren.BVH_AABB.itemsize == 48
    assert [f[1] for f in ren.struct_layout(ren.BVH_Triangle)] == [0, 16, 32, 48, 52]
    assert ren.BVH_Triangle.itemsize == 64
    assert ren.struct_layout(ren.Texture2D) == [("width", 0, 4), ("height", 4, 4), ("offset", 8, 4)]
    assert ren.Ray.itemsize == 32 and ren.Hit.itemsize == 20


def test_kernel_struct_alignment_rules():
    @ren.kernel_struct
    class S:
        a: np.int32
        b: ren.float3
        c: np.float32
        m: ren.float4x4

    assert [f[1] for f in ren.struct_layout(S)] == [0, 16, 32, 64]
    assert S.itemsize == 128


def test_buffers_map_get_and_clear_on_host():
    b = ren.create_buffer(10, ren.float3)
    assert b.shape == (10,) and b.dtype == ren.float3
    with ren.mapped(b) as m:
        m["x"] = np.arange(10)
    assert (b.get()["x"] == np.arange(10)).all()
    ren.clear(b, 2.5)
    assert (b.get()["y"] == 2.5).all()
    s = ren.create_struct(ren.Texture2D)
    with ren.mapped(s) as m:
        m["width"] = 7
    assert int(s.get()["width"]) == 7
    a = ren.create_buffer_from(np.arange(12, dtype=np.int32))
    np.testing.assert_array_equal(a.get(), np.arange(12))
    np.testing.assert_array_equal(a[3:6].get(), [3, 4, 5])


def test_host_camera_math_matches_closed_forms():
    eye, target, up = ren.make_float3(0, 0, 3), ren.make_float3(0, 0, 0), ren.make_float3(0, 1, 0)
    V = ren.to_array(ren.look_at(eye, target, up))
    # row-vector convention: world point @ V = view point; the target is straight ahead at z=3
    p = np.array([0, 0, 0, 1], np.float32) @ V
    np.testing.assert_allclose(p, [0, 0, 3, 1], atol=1e-6)
    # x axis is cross(up, z) = (-1, 0, 0) for z = (0, 0, -1): screen right is world -x
    np.testing.assert_allclose(np.array([-1, 0, 0, 0], np.float32) @ V, [1, 0, 0, 0], atol=1e-6)
    P = ren.to_array(ren.perspective(math.pi / 4, 2.0))
    hs = 1 / math.tan(math.pi / 8)
    assert abs(P[1, 1] - hs) < 1e-5 and abs(P[0, 0] - hs / 2) < 1e-5 and P[2, 3] == 1.0
    R = ren.to_array(ren.rotate(math.pi / 2, ren.make_float3(0, 0, 1)))
    np.testing.assert_allclose(np.array([1, 0, 0, 0], np.float32) @ R, [0, 1, 0, 0], atol=1e-6)
    T = ren.matmul(ren.translate(1, 2, 3), ren.scale(2.0))
    np.testing.assert_allclose(np.array([0, 0, 0, 1], np.float32) @ ren.to_array(T), [2, 4, 6, 1])


def test_camera_vectors_are_the_reference_basis():
    from rendertoy_amd._raycaster import Camera
    c = Camera((0, 0, 3), (0, 0, 0), (0, 1, 0), math.pi / 4, 16 / 9)
    t = math.tan(math.pi / 8)
    np.testing.assert_allclose(c.forward, [0, 0, -1], atol=1e-7)
    np.testing.assert_allclose(c.right, [-t * 16 / 9, 0, 0], rtol=1e-6)
    np.testing.assert_allclose(c.up, [0, t, 0], rtol=1e-6)


def test_dispatcher_launch_convention():
    calls = []
    d = ren.Dispatcher(lambda n, *a: calls.append((n, a)), "k")
    d[(4, 8)](1, 2)
    d[7]()
    assert calls == [(32, (1, 2)), (7, ())]


def test_mesh_triangle_count_rule():
    # rendering/_raycaster.py:33: V//3 without indices, I//3 with them
    m = ren.mesh_from_arrays(np.zeros((9, 3), np.float32))
    assert m.triangle_count == 3
    m2 = ren.mesh_from_arrays(np.zeros((4, 3), np.float32), np.array([[0, 1, 2], [0, 2, 3]]))
    assert m2.triangle_count == 2


def test_ray_casting_raises_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    m = ren.mesh_from_arrays(np.zeros((3, 3), np.float32))
    with pytest.raises(RuntimeError):
        ren.Raycaster([m])
