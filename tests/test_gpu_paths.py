"""Diffuse paths on the GPU (rt_trace_paths): radiance, ray counts per depth and the per-sample
Philox streams equal the CPU twin bit for bit, for every camera-ray traversal and for 0..3
bounces; a captured graph replays to the same image; an empty scene is all sky."""
import math

import numpy as np
import pytest
import torch

from helpers import twin_from_meshes, gpu_from_meshes, camera

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cluster(ren):
    from rendertoy_amd import scenes
    rng = np.random.default_rng(55)
    meshes = []
    for k in range(4):
        d = dict(scenes.icosphere(10 + 4 * k, 560 + k))
        d["pos"] = (d["pos"] * np.float32(0.5) + rng.uniform(-0.6, 0.6, 3).astype(np.float32)).astype(np.float32)
        meshes.append(d)
    return meshes, twin_from_meshes(meshes), gpu_from_meshes(meshes)


def run(rc, cam, W, H, **kw):
    b = kw.get("bounces", 2)
    counts = torch.zeros(b + 1, dtype=torch.int32, device="cuda")
    img = rc.trace_paths(cam, W, H, counts=counts, **kw)
    torch.cuda.synchronize()
    return img.cpu().numpy().reshape(-1), counts.cpu().numpy()


@pytest.mark.parametrize("primary", ["quad", "auto", "packet", "lane"])
@pytest.mark.parametrize("bounces,spp", [(2, 4), (0, 1), (3, 2), (2, 7)])
def test_paths_equal_twin(ren, cluster, primary, bounces, spp):
    meshes, tw, rc = cluster
    W, H = 96, 64
    cam = camera((0.4, 0.9, 2.6), (0, 0, 0), (0, 1, 0), math.pi / 4, W / H)
    with ren.tuning(primary=primary):
        img, cnt = run(rc, cam, W, H, spp=spp, bounces=bounces, frame=3)
    rad, smp, tcnt = tw.paths(cam.as_array(), W, H, spp=spp, bounces=bounces, frame=3)
    np.testing.assert_array_equal(img.view(np.uint32), rad.view(np.uint32))
    assert list(cnt) == list(tcnt)
    if bounces:
        assert tcnt[1] > 0


def test_paths_graph_replay_and_large_frame(ren, cluster):
    meshes, tw, rc = cluster
    W, H = 320, 200
    cam = camera((0.0, 0.2, 2.2), (0, 0, 0), (0, 1, 0), math.pi / 4, W / H)
    out = torch.empty((H, W), dtype=torch.float32, device="cuda")
    counts = torch.zeros(3, dtype=torch.int32, device="cuda")
    rc.trace_paths(cam, W, H, spp=4, bounces=2, out=out, counts=counts)
    torch.cuda.synchronize()
    ref = out.clone()
    rad, smp, tcnt = tw.paths(cam.as_array(), W, H, spp=4, bounces=2)
    np.testing.assert_array_equal(ref.cpu().numpy().reshape(-1).view(np.uint32), rad.view(np.uint32))
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        rc.trace_paths(cam, W, H, spp=4, bounces=2, out=out, counts=counts)
    for k in range(3):
        out.fill_(float("nan"))
        counts.fill_(-1)
        rc.render(cam, 64, 48)  # unrelated eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out.view(torch.int32), ref.view(torch.int32)), k
        assert list(counts.cpu().numpy()) == list(tcnt)


def test_paths_empty_scene_is_sky(ren):
    import rendertoy_amd as R
    rc = R.Raycaster([])
    cam = camera()
    img, cnt = run(rc, cam, 40, 30, spp=2, bounces=2, sky=0.5)
    assert (img == np.float32(0.5)).all() and list(cnt) == [40 * 30 * 2, 0, 0]
