"""A scene 16x the bench workload: the displaced icosphere of frequency 1024 (20,971,520 triangles,
10,485,762 vertices). Larger grids than the other tests exercise (radix-sort tiles, refit blocks,
sparse-table levels, top-list length): device == twin byte for byte for the LBVH build, the
camera hits, and the vertex normals against the generator's."""
import math
import time

import numpy as np
import pytest
import torch

from helpers import twin_from_meshes, camera, gpu_arrays

pytestmark = pytest.mark.gpu


def test_21m_triangle_build_trace_and_normals(ren):
    from rendertoy_amd import scenes
    t0 = time.time()
    d = scenes.icosphere(1024, 4000)
    print(f"generated {len(d['idx'])} triangles in {time.time() - t0:.1f} s", flush=True)
    mesh = ren.mesh_from_arrays(d["pos"], d["idx"], None, d["uv"])
    mesh.compute_normals()
    nrm = mesh.vertices.get().view(np.float32).reshape(-1, 20)[:, 4:7]
    np.testing.assert_array_equal(nrm.view(np.uint32), d["nrm"].view(np.uint32))
    del nrm
    print("normals equal", flush=True)
    rc = ren.Raycaster([mesh])
    torch.cuda.synchronize()
    tw = twin_from_meshes([{"pos": d["pos"], "idx": d["idx"], "nrm": d["nrm"], "uv": d["uv"]}])
    print(f"twin built, root {tw.root}", flush=True)
    assert rc.root == tw.root
    g = gpu_arrays(rc)
    np.testing.assert_array_equal(g["keys"].view(np.uint32), tw.keys)
    np.testing.assert_array_equal(g["order"].view(np.uint32), tw.order)
    for name in ("nodes4", "nodes", "tris", "shade", "bounds"):
        np.testing.assert_array_equal(np.ascontiguousarray(g[name]).view(np.uint32),
                                      np.ascontiguousarray(getattr(tw, name)).view(np.uint32))
        print(f"{name} equal", flush=True)
    del g
    W, H = 1920, 1080
    cam = camera(eye=(0.3, 0.2, 1.5), aspect=W / H)
    g = {k: v.cpu().numpy() for k, v in rc.cast_camera(cam, W, H).items()}
    t = tw.trace_camera(cam.as_array(), W, H, quad=True)
    np.testing.assert_array_equal(g["prim"], t["prim"])
    np.testing.assert_array_equal(g["t"].view(np.uint32), np.where(t["prim"] >= 0, t["t"], np.float32(np.inf)).view(np.uint32))
    print(f"camera hits equal ({int((g['prim'] >= 0).sum())} hits)", flush=True)
