"""Batched render (rt_render_batch): every job's image, hits and stats are bit-identical to the
job rendered alone with rt_render, for every launch shape; a captured graph replays to the same
result between unrelated eager work; the pixel-limit status word; the twin agrees with one job."""
import ctypes
import math

import numpy as np
import pytest
import torch

from helpers import gpu_from_meshes, twin_from_meshes, camera

pytestmark = pytest.mark.gpu

LIGHT = (np.array([1.0, 1.0, 1.0], np.float32) / np.float32(np.sqrt(3.0))).astype(np.float32)


@pytest.fixture(scope="module")
def scenes_(ren):
    from rendertoy_amd import scenes
    out = []
    for k, f in enumerate((3, 12, 40)):
        d = scenes.icosphere(f, 300 + k)
        M = scenes.random_transform(np.random.default_rng(400 + k))
        d = dict(d)
        d["pos"] = (d["pos"].astype(np.float64) @ M).astype(np.float32)
        nrm = d["nrm"].astype(np.float64) @ M
        d["nrm"] = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
        meshes = [d, scenes.ground_plane()] if k == 1 else [d]
        out.append((meshes, gpu_from_meshes(meshes)))
    tex = torch.from_numpy(scenes.value_noise_rgba8(64, 9)).cuda()
    torch.cuda.synchronize()
    return out, tex


def requests(ren, scenes_, with_hits=True):
    from rendertoy_amd.batch import RenderRequest
    sc, tex = scenes_
    specs = [  # (scene, W, H, eye, textured)
        (0, 64, 64, (0.0, 0.0, 3.0), False),
        (1, 200, 120, (0.5, 0.8, 2.5), True),
        (2, 37, 53, (-1.0, 0.3, 2.8), False),
        (1, 1, 1, (0.0, 0.0, 3.0), True),
        (0, 0, 16, (0.0, 0.0, 3.0), False),       # no pixels
        (2, 128, 96, (0.2, -0.4, 3.2), True),
        (0, 8, 8, (0.1, 0.1, 2.0), False),
        (1, 1024, 768, (0.3, 0.6, 2.6), True),   # coherent tiles: AUTO takes the packet walk
    ]
    reqs = []
    for s, W, H, eye, textured in specs:
        rc = sc[s][1]
        cam = camera(eye, (0, 0, 0), (0, 1, 0), math.pi / 4, (W / H) if H else 1.0)
        kw = {}
        if textured:
            kw = {"textures": [tex], "mesh_textures": [0] * len(sc[s][0])}
        reqs.append(RenderRequest(rc, cam, W, H, light=LIGHT, return_hits=with_hits, **kw))
    return reqs


def alone(r):
    stats = torch.zeros(4, dtype=torch.int32, device="cuda")
    img, hits = r.raycaster.render(r.camera, r.width, r.height, light=r.light, textures=r.textures,
                                   mesh_textures=r.mesh_textures, return_hits=True, stats=stats)
    return img.cpu().numpy(), {k: v.cpu().numpy() for k, v in hits.items()}, stats.cpu().numpy()


def assert_batch_equals_alone(b, reqs):
    for i, r in enumerate(reqs):
        img, hits, st = alone(r)
        n = r.width * r.height
        np.testing.assert_array_equal(b.images[i].cpu().numpy().view(np.uint32), img.view(np.uint32))
        np.testing.assert_array_equal(b.hits[i]["prim"].cpu().numpy()[:n], hits["prim"][:n])
        np.testing.assert_array_equal(b.hits[i]["t"].cpu().numpy()[:n].view(np.uint32),
                                      hits["t"][:n].view(np.uint32))
        np.testing.assert_array_equal(b.stats[i].cpu().numpy(), st)


@pytest.mark.parametrize("cfg", [dict(), dict(shadow="packet", primary="packet"),
                                 dict(tiles_per_grab=3, waves_per_cu=1, primary="lane"),
                                 dict(tiles_per_grab=64, waves_per_cu=32, lane_stack=2),
                                 dict(primary="hybrid", shadow="hybrid", packet_budget=2),
                                 dict(primary="packet", shadow="lane"), dict(primary="auto"),
                                 # per-lane refill kernel (LANE + LANE): stacks of 1 entry force the
                                 # packet retrace of both walks; grabs of 64 tiles cross jobs
                                 dict(primary="lane", shadow="lane", lane_stack=1, tiles_per_grab=2),
                                 dict(primary="lane", shadow="lane", tiles_per_grab=64, waves_per_cu=32),
                                 # quad-lane tiles (4x4 units) for both rays; grabs of 5 units cross jobs
                                 dict(primary="quad", tiles_per_grab=5, waves_per_cu=8),
                                 dict(primary="quad", tiles_per_grab=64, waves_per_cu=32)])
def test_batch_equals_individual_renders(ren, scenes_, cfg):
    from rendertoy_amd.batch import BatchRender, BatchConfig
    reqs = requests(ren, scenes_)
    b = BatchRender(reqs, config=BatchConfig(**cfg))
    b.launch()
    torch.cuda.synchronize()
    assert b.status() == 0
    assert_batch_equals_alone(b, reqs)


def test_batch_job_matches_twin(ren, scenes_):
    from rendertoy_amd.batch import BatchRender
    sc, tex = scenes_
    reqs = requests(ren, scenes_)[:1]
    b = BatchRender(reqs)
    b.launch()
    tw = twin_from_meshes(sc[0][0])
    r = reqs[0]
    rgb, t, prim, st = tw.render(r.camera.as_array(), r.width, r.height, LIGHT)
    np.testing.assert_array_equal(b.images[0].cpu().numpy().view(np.uint32), rgb.view(np.uint32))
    np.testing.assert_array_equal(b.hits[0]["prim"].cpu().numpy(), prim)
    assert list(b.stats[0].cpu().numpy()[:3]) == list(st)


def test_graph_replays_equal_eager(ren, scenes_):
    from rendertoy_amd.batch import BatchRender
    reqs = requests(ren, scenes_)
    b = BatchRender(reqs)
    b.launch()
    torch.cuda.synchronize()
    ref = [im.clone() for im in b.images]
    ref_stats = b.stats.clone()
    b.capture()
    other = requests(ren, scenes_, with_hits=False)[1]
    for k in range(3):
        for im in b.images:
            im.fill_(float("nan"))
        b.stats.fill_(-1)
        other.raycaster.render(other.camera, other.width, other.height)  # unrelated eager work
        torch.cuda.synchronize()
        b.replay()
        torch.cuda.synchronize()
        for im, rf in zip(b.images, ref):
            assert torch.equal(im.view(torch.int32), rf.view(torch.int32)), f"replay {k}"
        assert torch.equal(b.stats, ref_stats)


def test_unrenderable_jobs_set_status_and_are_skipped(ren, scenes_):
    """A job over RT_BATCH_MAX_JOB_PIXELS, or one without an rgb buffer, is not rendered (its
    buffers are never touched) and sets its status bit; the other jobs render normally."""
    from rendertoy_amd import _native as N
    from rendertoy_amd.batch import BatchRender
    reqs = requests(ren, scenes_)[:2]
    b = BatchRender(reqs)
    lib = N.load()
    for extra, want in (("oversized", 1), ("no_rgb", 2), ("both", 3)):
        jobs = (N.rt_render_job * 4).from_buffer_copy(
            bytes(b.jobs_dev.cpu().numpy().tobytes()) + bytes(2 * ctypes.sizeof(N.rt_render_job)))
        for k in (2, 3):
            jobs[k].scene = jobs[0].scene
            jobs[k].out = N.rt_render_out()                          # no output buffers at all
        if extra in ("oversized", "both"):
            jobs[2].params.width, jobs[2].params.height = 8192, 4096  # 2^25 pixels
        if extra in ("no_rgb", "both"):
            jobs[3].params.width, jobs[3].params.height = 16, 16     # rgb NULL
        jdev = torch.frombuffer(bytearray(bytes(jobs)), dtype=torch.uint8).cuda()
        ws = torch.zeros(lib.rt_render_batch_workspace_bytes(4), dtype=torch.uint8, device="cuda")
        cfg = b.config.to_c()
        for im in b.images:
            im.fill_(float("nan"))
        torch.cuda.synchronize()
        N.check(lib.rt_render_batch(jdev.data_ptr(), 4, ctypes.byref(cfg), ws.data_ptr(), ws.numel(),
                                    N.stream_handle()), "rt_render_batch")
        torch.cuda.synchronize()
        assert int(ws[4:8].view(torch.int32).item()) == want, extra
        assert_batch_equals_alone(b, reqs)


def test_relaunch_keeps_textures_the_caller_dropped(ren):
    """A batch holds every texture its jobs sample: the caller's own references are dropped before
    the first launch and the freed blocks are reused by other tensors, yet every relaunch renders
    the same images (the texels must not be freed under the job table)."""
    from rendertoy_amd import scenes
    from rendertoy_amd.batch import BatchRender, RenderRequest

    def make():
        rc = gpu_from_meshes([scenes.icosphere(10, 77), scenes.ground_plane()])
        tex = torch.from_numpy(scenes.value_noise_rgba8(32, 78)).cuda()
        ftex = torch.from_numpy(scenes.value_noise_rgba8(16, 79).astype(np.float32) / 255.0).cuda()
        reqs = []
        for k, t in enumerate((tex, ftex)):
            cam = camera((0.3 * k, 0.2, 2.6), (0, 0, 0), (0, 1, 0), math.pi / 4, 1.5)
            reqs.append(RenderRequest(rc, cam, 96, 64, light=LIGHT, textures=[t], mesh_textures=[0, 0]))
        return BatchRender(reqs)

    b = make()
    b.launch()
    torch.cuda.synchronize()
    want = [img.clone() for img in b.images]
    for seed in range(3):
        junk = [torch.full((32, 32, 4), 200 + seed, dtype=torch.uint8, device="cuda") for _ in range(64)]
        for img in b.images:
            img.fill_(-1.0)
        b.launch()
        torch.cuda.synchronize()
        del junk
        for w, g in zip(want, b.images):
            assert torch.equal(w.view(torch.int32), g.view(torch.int32))
    # the textured pixels are not all the albedo-1 gray
    assert any(bool((w[..., 0] != w[..., 1]).any()) for w in want)
