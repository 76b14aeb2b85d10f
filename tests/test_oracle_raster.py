"""Pins of the raster visibility oracle (tw_raster_ids: projection, near-plane clipping, screen-space
edge functions at pixel centres, nearest depth; the reference's raster rule of rendering/_raster.py
:80-112, :172-214, :282-311 re-expressed in FP64). CPU only.

- Against the FP64 ray oracle (closest64, a different method: rays and Moller-Trumbore in FP64) on
  scenes with occlusion, silhouettes and a ground plane that crosses the near plane: every pixel
  flagged by neither oracle sees the same triangle.
- Closed form: the pixels a screen-facing square covers follow from the camera's projection.
- Depth order does not depend on submission order; geometry behind the eye covers nothing."""
import math

import numpy as np

from helpers import camera
from rendertoy_amd import scenes

f32 = np.float32


def raster(pos, cam, W, H, **kw):
    import oracle
    return oracle.raster_ids(pos, cam.as_array(), W, H, **kw)


def fp64_rays(pos, cam, W, H):
    import oracle
    o, d = oracle.twin_generate_rays(cam.as_array(), W, H)
    return oracle.closest64(pos, o, d)


def test_raster_equals_fp64_rays_on_sphere_and_clipped_plane():
    d = scenes.icosphere(12, 77)
    g = scenes.ground_plane()
    pos = scenes.triangle_arrays([d, g])[0]
    W, H = 200, 150
    for eye in ((0.0, 0.0, 3.0), (1.5, 0.4, 1.2), (0.2, -0.45, 2.0)):
        cam = camera(eye=eye, aspect=W / H)
        r = raster(pos, cam, W, H)
        f = fp64_rays(pos, cam, W, H)
        free = (r["flags"] == 0) & (f["edge"] == 0)
        np.testing.assert_array_equal(r["prim"][free], f["prim"][free])
        assert free.mean() > 0.98
        assert (r["prim"] == len(pos) - 1).any() or (r["prim"] == len(pos) - 2).any()  # the plane is seen


def test_raster_equals_fp64_rays_on_random_overlapping_triangles():
    rng = np.random.default_rng(21)
    c = rng.uniform(-0.6, 0.6, (300, 1, 3))
    pos = (c + rng.normal(0, 0.15, (300, 3, 3))).astype(f32).reshape(-1, 9)
    W, H = 160, 120
    cam = camera(eye=(0.3, 0.2, 2.5), aspect=W / H)
    r = raster(pos, cam, W, H)
    f = fp64_rays(pos, cam, W, H)
    free = (r["flags"] == 0) & (f["edge"] == 0)
    np.testing.assert_array_equal(r["prim"][free], f["prim"][free])
    assert (r["prim"][free] >= 0).sum() > 5000


def test_square_coverage_closed_form():
    a = 0.5
    pos = np.array([[-a, -a, 0, a, -a, 0, a, a, 0], [-a, -a, 0, a, a, 0, -a, a, 0]], f32)
    W, H = 96, 64
    fov = math.pi / 4
    cam = camera(eye=(0.0, 0.0, 3.0), fov=fov, aspect=W / H)
    r = raster(pos, cam, W, H)
    th = math.tan(fov / 2)
    col, row = np.meshgrid(np.arange(W), np.arange(H))
    nx = (col + 0.5) / W * 2 - 1
    ny = 1 - (row + 0.5) / H * 2
    inside = (np.abs(nx) * 3 * th * (W / H) < a) & (np.abs(ny) * 3 * th < a)
    free = r["flags"].reshape(H, W) == 0
    np.testing.assert_array_equal((r["prim"].reshape(H, W) >= 0)[free], inside[free])
    assert inside.sum() > 300 and free.mean() > 0.9


def test_nearer_triangle_wins_in_any_order_and_behind_is_invisible():
    near = np.array([-1, -1, 0.5, 1, -1, 0.5, 0, 1, 0.5], f32)
    far = np.array([-2.5, -2.5, -0.5, 2.5, -2.5, -0.5, 0, 2.5, -0.5], f32)   # peeks out around near
    behind = np.array([-1, -1, 4.0, 1, -1, 4.0, 0, 1, 4.0], f32)   # the eye is at z = 3
    W, H = 64, 64
    cam = camera()
    r1 = raster(np.stack([near, far, behind]), cam, W, H)
    r2 = raster(np.stack([behind, far, near]), cam, W, H)
    m1 = np.array([0, 1, 2, -1])[r1["prim"]]              # map ids to (near, far, behind, none)
    m2 = np.array([2, 1, 0, -1])[r2["prim"]]
    np.testing.assert_array_equal(m1, m2)
    assert (m1 == 0).sum() > 200 and (m1 == 1).sum() > 50 and not (m1 == 2).any()
