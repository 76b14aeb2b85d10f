"""OBJ ingest (rendertoy_amd.load_obj / parse_obj; rendering/_loaders.py:7-42 without pywavefront).

CPU: face forms (v, v/vt, v//vn, v/vt/vn), negative indices, polygons split as fans from the first
corner, objects split at `o`, comments and ignored statements; the unit-cube normalisation is the
reference's scalar rule (worked by hand on a small box); per-corner vertex layout and the indexed
variant describe the same triangles. GPU: an icosphere written as OBJ and loaded back builds the same
scene and renders the same image as the arrays it was written from."""
import math

import numpy as np
import pytest

import rendertoy_amd as ren
from rendertoy_amd import _loaders, scenes

f32 = np.float32

OBJ = """# two objects
mtllib x.mtl
o quad
v 0 0 0
v 2 0 0
v 2 1 0
v 0 1 0
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vn 0 0 1
usemtl a
s off
f 1/1/1 2/2/1 3/3/1 4/4/1
o tri   # trailing comment
v 0 0 5
v 0 3 5
g grp
f -6//1 -1//1 -2//1
f 1 2 6
"""


def write_obj(path, pos, idx, nrm=None, uv=None):
    lines = [f"v {x!r} {y!r} {z!r}" for x, y, z in pos.astype(float).tolist()]
    if uv is not None:
        lines += [f"vt {u!r} {v!r}" for u, v in uv.astype(float).tolist()]
    if nrm is not None:
        lines += [f"vn {x!r} {y!r} {z!r}" for x, y, z in nrm.astype(float).tolist()]
    for a, b, c in (idx + 1).tolist():
        if uv is not None and nrm is not None:
            lines.append(f"f {a}/{a}/{a} {b}/{b}/{b} {c}/{c}/{c}")
        else:
            lines.append(f"f {a} {b} {c}")
    path.write_text("\n".join(lines) + "\n")


def test_parse_forms_fans_objects_and_negative_indices():
    obs = _loaders.parse_obj(OBJ)
    assert [o["name"] for o in obs] == ["quad", "tri"]
    q, t = obs
    # the quad is split as (1,2,3), (1,3,4)
    np.testing.assert_array_equal(q["pos"], f32([[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 0, 0], [2, 1, 0], [0, 1, 0]]))
    np.testing.assert_array_equal(q["uv"], f32([[0, 0], [1, 0], [1, 1], [0, 0], [1, 1], [0, 1]]))
    np.testing.assert_array_equal(q["nrm"], np.tile(f32([0, 0, 1]), (6, 1)))
    # -6 -> v1, -1 -> v6, -2 -> v5; the second face has no vt/vn
    np.testing.assert_array_equal(t["pos"], f32([[0, 0, 0], [0, 3, 5], [0, 0, 5], [0, 0, 0], [2, 0, 0], [0, 3, 5]]))
    assert t["uv"] is None
    np.testing.assert_array_equal(t["nrm"], f32([[0, 0, 1]] * 3 + [[0, 0, 0]] * 3))


def test_parse_errors():
    with pytest.raises(ValueError):
        _loaders.parse_obj("v 0 0 0\nv 1 0 0\nf 1 2 3\n")        # index out of range
    with pytest.raises(ValueError):
        _loaders.parse_obj("v 0 0 0\nv 1 0 0\nf 1 2\n")          # two corners


def test_unit_cube_rule_by_hand():
    # x in [1, 3], y in [0, 1], z in [-1, 0]: scalar min -1, max 3, size 4
    p = f32([[1, 0, -1], [3, 1, 0]])
    n = _loaders.normalize_unit_cube(p)
    np.testing.assert_array_equal(n, f32([[0.0, -0.25, -0.5], [0.5, 0.0, -0.25]]))


def test_load_obj_layouts(tmp_path):
    path = tmp_path / "two.obj"
    path.write_text(OBJ)
    objs = ren.load_obj(str(path))
    assert len(objs) == 2 and all(mat is None for _, mat in objs)
    (mq, _), (mt, _) = objs
    assert mq.indices is None and mq.vertices.shape == (6,) and mq.triangle_count == 2
    v = mq.vertices.get().view(f32).reshape(6, 20)
    want = _loaders.normalize_unit_cube(_loaders.parse_obj(OBJ)[0]["pos"])
    np.testing.assert_array_equal(v[:, 0:3], want)
    np.testing.assert_array_equal(v[:, 8:10], _loaders.parse_obj(OBJ)[0]["uv"])
    assert mq.has_normals and mq.has_uvs and mt.has_normals and not mt.has_uvs
    # indexed: 4 distinct corners for the quad, same triangles
    (iq, _), _ = ren.load_obj(str(path), indexed=True)
    iv = iq.vertices.get().view(f32).reshape(-1, 20)
    ii = iq.indices.get().reshape(-1, 3)
    assert iv.shape[0] == 4
    np.testing.assert_array_equal(iv[ii.reshape(-1), 0:3], want)


@pytest.mark.gpu
def test_obj_round_trip_builds_and_renders_the_same(ren, tmp_path):
    d = scenes.icosphere(24, 55)
    path = tmp_path / "s.obj"
    write_obj(path, d["pos"], d["idx"], d["nrm"], d["uv"])
    (m, _), = ren.load_obj(str(path))
    idx = d["idx"].reshape(-1)
    direct = ren.mesh_from_arrays(_loaders.normalize_unit_cube(d["pos"][idx]), None, d["nrm"][idx], d["uv"][idx])
    a, b = ren.Raycaster([m]), ren.Raycaster([direct])
    for k in ("_nodes4", "_tris", "_shade"):
        assert bool((getattr(a, k) == getattr(b, k)).all()), k
    W, H = 128, 96
    cam = ren.Camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.pi / 4, W / H)
    light = (np.array([1, 1, 1], f32) / f32(np.sqrt(3.0))).astype(f32)
    ia = a.render(cam, W, H, light=light).cpu().numpy()
    ib = b.render(cam, W, H, light=light).cpu().numpy()
    np.testing.assert_array_equal(ia.view(np.uint32), ib.view(np.uint32))
    # the indexed load shares vertices, so device normals are smooth and bit-equal to numpy's
    (mi, _), = ren.load_obj(str(path), indexed=True)
    mi.has_normals = False
    mi.compute_normals()
    pos_i = mi.vertices.get().view(f32).reshape(-1, 20)[:, 0:3]
    idx_i = mi.indices.get().reshape(-1, 3)
    np.testing.assert_array_equal(mi.vertices.get().view(f32).reshape(-1, 20)[:, 4:7].view(np.uint32),
                                  scenes.vertex_normals(pos_i, idx_i).view(np.uint32))
