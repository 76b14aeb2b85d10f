"""The reference-side binding printed in INTEGRATION.md, run as written (the block is read out of
the file, only the library path is filled in): its ctypes structs equal the C header's layouts
(CPU), and its Raycaster, built from Mesh objects, returns exactly the hits of the drop-in's
Raycaster.ray_cast (GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MARK = "# rendering/_raycaster.py (reference), additions"


def binding_namespace():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```python\n(.*?)```", text, flags=re.S)
    code = next(b for b in blocks if b.startswith(MARK))
    from rendertoy_amd import _native
    _native.load()  # torch first, then the library: one HIP runtime
    code = code.replace('ctypes.CDLL("librtx_hip.so")', f"ctypes.CDLL({_native.LIB_PATH!r})")
    ns = {"__name__": "reference_binding"}
    exec(compile(code, "INTEGRATION.md", "exec"), ns)
    return ns


def test_binding_structs_match_the_library_structs():
    from rendertoy_amd import _native as N
    ns = binding_namespace()
    for mine, theirs in (("_Scene", N.rt_scene), ("_Sizes", N.rt_scene_sizes), ("_MeshDesc", N.rt_mesh_desc),
                         ("_Rays", N.rt_rays), ("_Hits", N.rt_hits)):
        a = ns[mine]
        assert ctypes.sizeof(a) == ctypes.sizeof(theirs), mine
        for (name, _), (name2, _) in zip(a._fields_, theirs._fields_):
            assert name == name2 and getattr(a, name).offset == getattr(theirs, name2).offset, (mine, name)
    import rendertoy_amd as ren
    assert ns["HIT"].itemsize == ren.Hit.itemsize
    assert [ns["HIT"].fields[k][1] for k in ns["HIT"].names] == [ren.Hit.fields[k][1] for k in ren.Hit.names]


@pytest.mark.gpu
def test_binding_ray_cast_equals_the_drop_in(ren):
    from helpers import random_rays
    from rendertoy_amd import scenes
    ns = binding_namespace()
    a = scenes.icosphere(20, 211)
    b = scenes.icosphere(4, 212)
    b["pos"] = (b["pos"] * 0.3 + np.float32(0.45)).astype(np.float32)
    models = [ren.mesh_from_arrays(a["pos"], a["idx"], a["nrm"], a["uv"]), ren.mesh_from_arrays(b["pos"], b["idx"])]
    o, d = random_rays(4000, 213)
    rec = np.zeros(len(o), ren.Ray)
    v = rec.view(np.float32).reshape(-1, 8)
    v[:, 0:3], v[:, 4:7] = o, d
    rays = ren.create_buffer_from(rec)
    theirs = ns["Raycaster"](models).ray_cast(rays)
    mine = ren.Raycaster(models).ray_cast(rays).get()
    assert theirs.dtype.names == mine.dtype.names
    np.testing.assert_array_equal(theirs.view(np.uint8), mine.view(np.uint8))
    assert (theirs["index"] >= 0).sum() > 100 and (theirs["mesh"] == 1).sum() > 10
