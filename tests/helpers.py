"""Shared helpers of the test suite: build the same scene on the GPU (product) and in the twin
(oracle), and compare."""
import math

import numpy as np


def twin_from_meshes(meshes, builder="lbvh"):
    import oracle
    from rendertoy_amd import scenes
    pos, nrm, uv, mo = scenes.triangle_arrays(meshes)
    has_n = [m.get("nrm") is not None for m in meshes]
    has_uv = [m.get("uv") is not None for m in meshes]
    return oracle.TwinScene(pos, nrm, uv, mo, has_n, has_uv, builder=builder)


def gpu_from_meshes(meshes, indexed=True):
    import rendertoy_amd as ren
    models = []
    for m in meshes:
        if indexed:
            mm = ren.mesh_from_arrays(m["pos"], m["idx"], m.get("nrm"), m.get("uv"))
        else:
            idx = m["idx"]
            mm = ren.mesh_from_arrays(m["pos"][idx].reshape(-1, 3), None,
                                      None if m.get("nrm") is None else m["nrm"][idx].reshape(-1, 3),
                                      None if m.get("uv") is None else m["uv"][idx].reshape(-1, 2))
        models.append(mm)
    return ren.Raycaster(models)


def gpu_arrays(rc):
    a = rc.scene_arrays()
    return {k: v.cpu().numpy() for k, v in a.items()}


def camera(eye=(0.0, 0.0, 3.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=math.pi / 4,
           aspect=1.0):
    from rendertoy_amd._raycaster import Camera
    return Camera(eye, target, up, fov, aspect)


def random_rays(n, seed, box=1.5):
    """Rays from random origins on a sphere of radius 3 toward random points in the unit cube."""
    rng = np.random.default_rng(seed)
    o = rng.standard_normal((n, 3))
    o = 3.0 * o / np.linalg.norm(o, axis=1, keepdims=True)
    tgt = rng.uniform(-0.5 * box, 0.5 * box, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(np.float32), d.astype(np.float32)


def soa_tensors(o, d):
    import torch
    return {k: torch.from_numpy(np.ascontiguousarray(a[:, i])).cuda()
            for k, a, i in (("ox", o, 0), ("oy", o, 1), ("oz", o, 2),
                            ("dx", d, 0), ("dy", d, 1), ("dz", d, 2))}
