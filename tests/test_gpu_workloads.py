"""The batched, serving and incoherent-ray workloads run at their full size on the GPU:

- batched: the 256-scene batch (1k-500k triangles each, 128^2-1024^2 images, seeds 3000 + i) in
  one graph-captured persistent launch; every job bit-equal to its render alone, four jobs equal
  to the twin;
- serving: the 64-scene pool (10k-2M triangles, 512^2 textures) behind a budget that forces
  evictions and rebuilds; 96 requests of 256^2-1024^2 served equal to their renders alone, two
  equal to the twin;
- incoherent: the 8 x 250,880-triangle cluster at 1920x1080 x 4 spp x 2 bounces equal to the twin
  (radiance and rays per depth), from the bench's view and from a view that frames the cluster.
"""
import math

import numpy as np
import pytest
import torch

from helpers import camera, twin_from_meshes

pytestmark = pytest.mark.gpu

LIGHT = (np.array([1.0, 1.0, 1.0], np.float32) / np.float32(np.sqrt(3.0))).astype(np.float32)


def _u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_batched_workload_256_scenes(ren, twin):
    from rendertoy_amd import scenes
    from rendertoy_amd.batch import BatchRender, RenderRequest
    reqs, meshes, tris = [], [], 0
    for i in range(256):
        d, res = scenes.batch_scene(i)
        rc = ren.Raycaster([ren.mesh_from_arrays(d["pos"], d["idx"], d["nrm"], d["uv"])])
        cam = camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.pi / 4, 1.0)
        reqs.append(RenderRequest(rc, cam, res, res, light=LIGHT, return_hits=True))
        meshes.append(d)
        tris += rc.n_tris
    print(f"batched: {len(reqs)} scenes, {tris} triangles, {sum(r.width * r.height for r in reqs)} pixels")
    assert 256 == len(reqs) and tris > 10_000_000
    b = BatchRender(reqs)
    b.capture()
    b.replay()
    torch.cuda.synchronize()
    assert b.status() == 0
    for i, r in enumerate(reqs):
        stats = torch.zeros(4, dtype=torch.int32, device="cuda")
        img, hits = r.raycaster.render(r.camera, r.width, r.height, light=LIGHT, return_hits=True, stats=stats)
        assert torch.equal(b.images[i].view(torch.int32), img.view(torch.int32)), i
        assert torch.equal(b.hits[i]["prim"], hits["prim"]), i
        assert torch.equal(b.stats[i], stats), i
    # four jobs against the twin: the smallest and largest scenes at two image sizes
    order = np.argsort([len(m["idx"]) for m in meshes])
    for i in (int(order[0]), int(order[1]), int(order[-1]), int(order[len(order) // 2])):
        r = reqs[i]
        tw = twin_from_meshes([meshes[i]])
        rgb, t, prim, st = tw.render(r.camera.as_array(), r.width, r.height, LIGHT)
        np.testing.assert_array_equal(_u32(b.images[i].cpu().numpy()), _u32(rgb))
        np.testing.assert_array_equal(b.hits[i]["prim"].cpu().numpy(), prim)
        assert list(b.stats[i].cpu().numpy()[:3]) == list(st)


def test_serving_workload_64_scene_pool_under_eviction(ren, twin):
    from rendertoy_amd import scenes, serving
    specs = {}

    def spec(k):
        d, tex = scenes.serving_scene(k)
        return serving.SceneSpec([d], tex)
    for k in range(64):
        specs[k] = (lambda k=k: spec(k))
    budget = 2 << 30  # 2 GiB: about a third of the pool (5.9 GB) fits
    store = serving.SceneStore(specs, budget)
    reqs = serving.make_requests(96, 64, 2000.0, 4999)
    for r in reqs:
        r.keep_image = True
    srv = serving.RenderServer(store, max_batch=16)
    srv.run(reqs, realtime=False)
    torch.cuda.synchronize()
    print(f"serving: {len(reqs)} requests in {len(srv.batches)} batches, store hits {store.hits} "
          f"misses {store.misses} evictions {store.evictions}")
    assert store.evictions > 0 and sum(srv.batches) == len(reqs)
    assert {r.size for r in reqs} == {256, 512, 1024}
    for r in reqs:
        res = store.acquire(r.scene)
        img = res.raycaster.render(serving.camera_for(r.eye), r.size, r.size, textures=res.binding)
        store.release(res)
        assert torch.equal(img.view(torch.int32), r.image.view(torch.int32)), (r.scene, r.size)
    # two served requests against the twin (the smallest scene among them, and a 1024^2 one)
    small = min(reqs, key=lambda r: specs[r.scene]().tris())
    big = next(r for r in reqs if r.size == 1024)
    for r in (small, big):
        sp = specs[r.scene]()
        tw = twin_from_meshes(sp.meshes)
        cam = serving.camera_for(r.eye)
        rgb, _, _, _ = tw.render(cam.as_array(), r.size, r.size, LIGHT, textures=[sp.texture], mesh_tex=[0])
        np.testing.assert_array_equal(_u32(r.image.cpu().numpy()), _u32(rgb))


@pytest.mark.parametrize("view", ["bench", "framed"])
def test_incoherent_workload_full_frame(ren, twin, view):
    from rendertoy_amd import scenes
    meshes = scenes.cluster_meshes()
    rc = ren.Raycaster([ren.mesh_from_arrays(m["pos"], m["idx"], m["nrm"], m["uv"]) for m in meshes])
    tw = twin_from_meshes(meshes)
    W, H, spp, bounces = 1920, 1080, 4, 2
    eye, fov, target = ((0.0, 0.0, 3.0), math.pi / 4, (0.0, 0.0, 0.0)) if view == "bench" \
        else scenes.cluster_view()
    cam = camera(eye, target, (0, 1, 0), fov, W / H)
    counts = torch.zeros(bounces + 1, dtype=torch.int32, device="cuda")
    img = rc.trace_paths(cam, W, H, spp=spp, bounces=bounces, frame=3, counts=counts)
    rad, _, cnt = tw.paths(cam.as_array(), W, H, spp=spp, bounces=bounces, frame=3)
    print(f"incoherent ({view}): rays per depth {counts.cpu().numpy().tolist()}")
    np.testing.assert_array_equal(_u32(img.cpu().numpy().ravel()), _u32(rad))
    np.testing.assert_array_equal(counts.cpu().numpy(), cnt)
    if view == "framed":
        assert cnt.sum() > 0.6 * W * H * spp * (bounces + 1)  # most samples hit and bounce
