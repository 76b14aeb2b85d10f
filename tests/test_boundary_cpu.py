"""Host-side pieces of the boundary that need no GPU: the camera read from the reference's
look_at / perspective matrices, the index narrowing rule, the launch-count and ray-field checks,
and the per-call tuning records."""
import math

import numpy as np
import pytest
import torch

import rendertoy_amd as ren
from rendertoy_amd import _raycaster as R


@pytest.mark.parametrize("eye,target,up,fov,aspect", [
    ((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.pi / 4, 1.0),
    ((0.7, 0.9, 2.1), (0.05, -0.1, 0.0), (0.0, 1.0, 0.0), 0.9, 199 / 131),
    ((-2.0, 1.0, -1.5), (0.3, 0.2, 0.1), (0.0, 0.0, 1.0), 1.3, 0.5),
])
def test_camera_from_matrices_matches_the_direct_camera(eye, target, up, fov, aspect):
    view = ren.look_at(ren.make_float3(*eye), ren.make_float3(*target), ren.make_float3(*up))
    proj = ren.perspective(fov, aspect)
    a = R.Camera.from_matrices(view, proj)
    b = R.Camera(eye, target, up, fov, aspect)
    for k in ("eye", "right", "up", "forward"):
        np.testing.assert_allclose(getattr(a, k), getattr(b, k), rtol=0, atol=3e-6)
    assert abs(a.fov - fov) < 1e-6 and abs(a.aspect - aspect) < 1e-5
    # the basis is the view matrix's columns, bit for bit
    V = ren.to_array(view).reshape(4, 4)
    np.testing.assert_array_equal(a.forward, V[:3, 2])
    # a point on the view axis projects to the image centre: forward is the centre ray
    M = ren.to_array(ren.matmul(ren.make_float4(*(np.asarray(eye) + 2.0 * a.forward), 1.0), view))
    np.testing.assert_allclose(M[:2], 0.0, atol=1e-5)


def test_camera_from_matrices_rejects_other_projections():
    view = ren.look_at(ren.make_float3(0, 0, 3), ren.make_float3(0, 0, 0), ren.make_float3(0, 1, 0))
    with pytest.raises(ValueError, match="perspective"):
        R.Camera.from_matrices(view, ren.identity())


def test_index_narrowing_clamps_out_of_range_values():
    idx = np.array([0, 5, 49, 50, -1, 2 ** 32 + 5, -(2 ** 33)], np.int64)
    buf = ren.create_buffer_from(idx)
    got = R._narrow_indices(buf, 50).cpu().numpy()
    np.testing.assert_array_equal(got, [0, 5, 49, 49, 49, 49, 49])
    u16 = ren.create_buffer_from(np.array([0, 40000, 65535, 7], np.uint16))
    np.testing.assert_array_equal(R._narrow_indices(u16, 70000).cpu().numpy(), [0, 40000, 65535, 7])
    np.testing.assert_array_equal(R._narrow_indices(u16, 100).cpu().numpy(), [0, 99, 99, 7])
    i16 = ren.create_buffer_from(np.array([-3, 12], np.int16))
    np.testing.assert_array_equal(R._narrow_indices(i16, 20).cpu().numpy(), [19, 12])
    with pytest.raises(TypeError, match="index buffers"):
        R._narrow_indices(ren.create_buffer_from(np.zeros(3, np.float32)), 10)


def bare_raycaster():
    """A Raycaster with no scene (no device work): enough for the argument checks."""
    return R.Raycaster.__new__(R.Raycaster)


def test_ray_field_checks():
    rc = bare_raycaster()
    n = 10
    rays = {k: torch.zeros(n, dtype=torch.float32) for k in ("ox", "oy", "oz", "dx", "dy", "dz")}
    with pytest.raises(RuntimeError, match="HIP device"):
        rc._rays_desc(rays, 0.0, math.inf)
    with pytest.raises(KeyError):
        rc._rays_desc({"ox": rays["ox"]}, 0.0, math.inf)
    with pytest.raises(TypeError):
        rc._rays_desc([1, 2, 3], 0.0, math.inf)


def test_dispatcher_launch_count():
    seen = []
    d = ren.Dispatcher(lambda n, *a: seen.append((n, a)))
    d[7]("x")
    d[(3, 4)]("y", 1)
    assert seen == [(7, ("x",)), (12, ("y", 1))]


def test_make_tuning_merges_fields():
    from rendertoy_amd.tuning import make_tuning, as_dict, LANE, PLOC
    base = ren.get_tuning()
    t = make_tuning(None, query="lane", lane_stack=8, builder="ploc")
    d = as_dict(t)
    assert d["query"] == LANE and d["lane_stack"] == 8 and d["builder"] == PLOC
    assert {k: v for k, v in d.items() if k not in ("query", "lane_stack", "builder")} == \
        {k: v for k, v in base.items() if k not in ("query", "lane_stack", "builder")}
    with pytest.raises(TypeError):
        make_tuning(None, bogus=1)
