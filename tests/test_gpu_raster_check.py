"""Primary visibility of the device (rt_trace_camera) against the raster oracle (tw_raster_ids):
a second statement of which triangle each pixel centre sees, reached by projection, near-plane
clipping, screen-space edge functions and a depth buffer instead of rays (the reference's raster
rule, rendering/_raster.py:80-112, :172-214, :282-311, re-expressed in FP64). Every pixel the
raster does not flag (a triangle edge within 2e-3 px of the centre, or another surface within the
depth range that a 2e-3 px shift spans) must see the same triangle."""
import math

import numpy as np
import pytest
import torch

from conftest import c1_scene, c2_scene
from helpers import gpu_from_meshes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["c1", "c2", "c2_side"])
def test_primary_ids_match_the_raster(ren, case):
    import oracle
    if case == "c1":
        meshes, W, H, eye = c1_scene(), 512, 512, (0.0, 0.0, 3.0)
    else:
        meshes, W, H = c2_scene()[0], 1920, 1080
        eye = (0.0, 0.0, 3.0) if case == "c2" else (2.2, 0.3, 1.6)
    rc = gpu_from_meshes(meshes)
    cam = ren.Camera(eye, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.pi / 4, W / H)
    out = rc.cast_camera(cam, W, H)
    prim = out["prim"].cpu().numpy()
    from rendertoy_amd import scenes
    pos = scenes.triangle_arrays(meshes)[0]
    r = oracle.raster_ids(pos, cam.as_array(), W, H)
    mism = prim != r["prim"]
    free = r["flags"] == 0
    print(f"{case}: pixels={W * H} hits={int((prim >= 0).sum())} raster_flagged={int((~free).sum())} "
          f"mismatches={int(mism.sum())} unflagged_mismatches={int((mism & free).sum())}", flush=True)
    assert int((mism & free).sum()) == 0
    assert free.mean() > 0.99
