"""The allocation-free batch entry point (rt_render_batch_static), called through the C ABI: two
calls in a row, the second with fewer jobs, each bit-identical to rt_render_batch on the same job
table (which uses a workspace argument); and its argument checks."""
import ctypes
import math

import numpy as np
import pytest
import torch

from helpers import gpu_from_meshes, camera

pytestmark = pytest.mark.gpu

LIGHT = (np.array([1.0, 1.0, 1.0], np.float32) / np.float32(np.sqrt(3.0))).astype(np.float32)


def batch(ren, n_jobs, seed):
    from rendertoy_amd import scenes
    from rendertoy_amd.batch import BatchRender, RenderRequest
    rng = np.random.default_rng(seed)
    a = gpu_from_meshes([scenes.icosphere(14, seed), scenes.ground_plane()])
    b = gpu_from_meshes([scenes.icosphere(5, seed + 1)])
    tex = torch.from_numpy(scenes.value_noise_rgba8(32, seed + 2)).cuda()
    reqs = []
    for k in range(n_jobs):
        rc = a if k % 2 == 0 else b
        W, H = int(rng.integers(1, 200)), int(rng.integers(1, 150))
        eye = tuple(rng.uniform(-0.5, 0.5, 2)) + (2.5,)
        kw = {"textures": [tex], "mesh_textures": [0, -1]} if rc is a else {}
        reqs.append(RenderRequest(rc, camera(eye, (0, 0, 0), (0, 1, 0), math.pi / 4, W / H), W, H,
                                  light=LIGHT, return_hits=True, **kw))
    return BatchRender(reqs)


def outputs(b):
    return [(img.view(torch.int32).clone(), h["prim"].clone(), h["t"].view(torch.int32).clone())
            for img, h in zip(b.images, b.hits)]


def test_static_batch_equals_the_workspace_batch(ren):
    from rendertoy_amd import _native as N
    lib = N.load()
    for n_jobs, seed in ((40, 901), (7, 902)):   # second call: a smaller batch on the same scratch
        b = batch(ren, n_jobs, seed)
        b.launch()
        torch.cuda.synchronize()
        want = outputs(b)
        for img, h in zip(b.images, b.hits):
            img.fill_(-1.0)
            h["prim"].fill_(-7)
        cfg = b.config.to_c()
        N.check(lib.rt_render_batch_static(b.jobs_dev.data_ptr(), b.n_jobs, ctypes.byref(cfg),
                                           N.stream_handle()), "rt_render_batch_static")
        torch.cuda.synchronize()
        got = outputs(b)
        for (wi, wp, wt), (gi, gp, gt) in zip(want, got):
            assert torch.equal(wi, gi) and torch.equal(wp, gp) and torch.equal(wt, gt)


def test_static_batch_argument_checks(ren):
    from rendertoy_amd import _native as N
    lib = N.load()
    cfg = N.rt_batch_config(8, 1, N.RT_TRAVERSAL_QUAD, N.RT_TRAVERSAL_LANE, 32, 256)
    assert lib.rt_render_batch_static(None, 1, ctypes.byref(cfg), None) == -1
    assert lib.rt_render_batch_static(0x1000, N.RT_BATCH_MAX_JOBS + 1, ctypes.byref(cfg), None) == -3
    assert b"RT_BATCH_MAX_JOBS" in lib.rt_last_error()
    bad = N.rt_batch_config(0, 1, N.RT_TRAVERSAL_QUAD, N.RT_TRAVERSAL_LANE, 32, 256)
    assert lib.rt_render_batch_static(0x1000, 1, ctypes.byref(bad), None) == -1
    assert lib.rt_render_batch_static(0x1000, 0, ctypes.byref(cfg), None) == 0
