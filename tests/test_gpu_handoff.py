"""The camera kernel's hand-off of its slowest packets (rt_trace_camera with a workspace): packets
past handoff_steps node steps queue their remaining subtrees, which every wave of the GPU walks
(re-queued in pieces after item_steps steps) and merges by a 64-bit atomic minimum. The hits must be
bit-identical to the walk without a workspace and to the twin, for the default budgets, for budgets
of one step (nearly every packet handed off, every item split again), and when the queue fills."""
import math

import numpy as np
import pytest
import torch

from conftest import c2_scene
from helpers import camera, gpu_from_meshes, twin_from_meshes

# The hand-off is switched off (rt_tuning.handoff_steps must be 0): the first frame with it at
# 1920x1080 (handoff 48, item 32) did not finish, both before and after its queue's waits were made
# relaxed and bounded by a deadline, and no cause has been shown. These tests stay for when it is.
pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="camera hand-off switched off: its 1920x1080 run did not finish, "
                                      "cause not shown")]


def _u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


def cast(rc, cam, W, H, **tuning):
    out = rc.cast_camera(cam, W, H, fields=("t", "u", "v", "prim", "index", "mesh"), tuning=tuning or None)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def status_words(rc):
    return [int(ws[:4].view(torch.int32).item()) for ws in rc._cam_ws.values()]


def assert_same(a, b):
    for k in ("prim", "index", "mesh"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    for k in ("t", "u", "v"):
        np.testing.assert_array_equal(_u32(a[k]), _u32(b[k]), err_msg=k)


@pytest.fixture(scope="module")
def c2(ren):
    meshes, tex = c2_scene()
    rc = gpu_from_meshes(meshes)
    tw = twin_from_meshes(meshes)
    torch.cuda.synchronize()
    return meshes, rc, tw


def test_c2_frame_with_handoff_equals_twin(c2, twin):
    meshes, rc, tw = c2
    W, H = 1920, 1080
    cam = camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.pi / 4, W / H)
    g = cast(rc, cam, W, H, handoff_steps=48, item_steps=32)
    assert rc._cam_ws, "the camera path takes a workspace"
    ref = tw.trace_camera(cam.as_array(), W, H, quad=True)
    np.testing.assert_array_equal(g["prim"], ref["prim"])
    for k in ("t", "u", "v"):
        np.testing.assert_array_equal(_u32(g[k]), _u32(ref[k]))
    plain = cast(rc, cam, W, H, handoff_steps=0)
    assert_same(g, plain)
    assert status_words(rc) == [0] * len(rc._cam_ws)


@pytest.mark.parametrize("handoff,item", [(1, 1), (2, 1), (8, 3), (48, 32), (200, 500)])
def test_tiny_budgets_give_identical_hits(c2, handoff, item):
    meshes, rc, tw = c2
    W, H = 640, 360
    for eye in ((0.0, 0.0, 3.0), (1.2, -0.3, 1.4)):
        cam = camera(eye, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.pi / 4, W / H)
        want = cast(rc, cam, W, H, handoff_steps=0)
        got = cast(rc, cam, W, H, handoff_steps=handoff, item_steps=item)
        assert_same(got, want)
        assert status_words(rc) == [0] * len(rc._cam_ws)


def test_full_queue_falls_back_to_walking_on(c2):
    """Budgets of one step at 1920x1080 queue more items than the queue holds (8 per packet): the
    packets that find it full walk on in their own wave, with the same hits."""
    meshes, rc, tw = c2
    W, H = 1920, 1080
    cam = camera((0.4, -0.35, 1.1), (0.0, -0.2, 0.0), (0.0, 1.0, 0.0), 1.0, W / H)
    want = cast(rc, cam, W, H, handoff_steps=0)
    got = cast(rc, cam, W, H, handoff_steps=1, item_steps=1)
    assert_same(got, want)
    assert status_words(rc) == [0] * len(rc._cam_ws)


def test_repeated_calls_and_odd_sizes(ren, twin):
    from rendertoy_amd import scenes
    meshes = [scenes.icosphere(40, 77), scenes.ground_plane()]
    rc = gpu_from_meshes(meshes)
    tw = twin_from_meshes(meshes)
    for W, H in ((1, 1), (3, 5), (37, 23), (501, 333), (64, 64)):
        cam = camera((0.2, 0.1, 2.2), (0.0, -0.1, 0.0), (0.0, 1.0, 0.0), 0.9, W / H)
        for rep in range(2):
            g = cast(rc, cam, W, H, handoff_steps=2, item_steps=2)
            ref = tw.trace_camera(cam.as_array(), W, H, quad=True)
            np.testing.assert_array_equal(g["prim"], ref["prim"])
            np.testing.assert_array_equal(_u32(g["t"]), _u32(ref["t"]))


def test_empty_and_single_leaf_scenes(ren):
    rng = np.random.default_rng(5)
    for n in (0, 1, 3):
        pos = rng.uniform(-0.5, 0.5, (max(3 * n, 3), 3)).astype(np.float32)
        idx = np.arange(3 * n, dtype=np.int32).reshape(-1, 3)
        rc = ren.Raycaster([ren.mesh_from_arrays(pos, idx)])
        cam = camera((0.0, 0.0, 2.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 1.2, 1.0)
        a = cast(rc, cam, 33, 33, handoff_steps=1, item_steps=1)
        b = cast(rc, cam, 33, 33, handoff_steps=0)
        assert_same(a, b)
