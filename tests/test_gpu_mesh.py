"""GPU parity of the mesh entry points of rt_mesh.hip.

- Per-vertex normals (rt_mesh_compute_normals via Mesh.compute_normals and
  rendertoy_amd.compute_normals) against the oracle's twin, bit for bit, across the sort's pass
  counts (1, 2, 3 and 4 passes of 8 bits over the vertex ids), index kinds and edge cases.
- MeshVertex records <-> dense tensors (rt_mesh_pack / rt_mesh_unpack) against the host packing
  of mesh_from_arrays, byte for byte, and round trips."""
import numpy as np
import pytest
import torch

from conftest import c2_scene
from helpers import camera

pytestmark = pytest.mark.gpu

f32 = np.float32


def twin_normals(pos, idx):
    import oracle
    return oracle.vertex_normals(pos, idx)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def record_floats(mesh):
    v = mesh.vertices.get()
    return v.view(np.float32).reshape(len(v), v.itemsize // 4)


def check(ren, pos, idx, int64=False):
    """In place (N field of the records) and into a dense tensor; both equal the twin."""
    m = ren.mesh_from_arrays(pos, idx)
    if int64 and idx is not None:
        m = ren.Mesh(m.vertices, ren.create_buffer_from(np.asarray(idx, np.int64).reshape(-1)))
        m.has_normals, m.has_uvs = False, False
    before = record_floats(m).copy()
    out = torch.empty((len(pos), 3), dtype=torch.float32, device="cuda")
    ren.compute_normals(m, out=out)
    m.compute_normals()
    after = record_floats(m)
    want = twin_normals(pos, idx)
    print(f"normals: V={len(pos)} T={0 if idx is None else len(np.asarray(idx).reshape(-1, 3))} "
          f"indexed={idx is not None} int64={int64}", flush=True)
    np.testing.assert_array_equal(bits(after[:, 4:7]), bits(want))
    np.testing.assert_array_equal(bits(out.cpu().numpy()), bits(want))
    keep = np.ones(after.shape[1], bool)
    keep[4:7] = False
    np.testing.assert_array_equal(bits(after[:, keep]), bits(before[:, keep]))  # P, pad, C, T, B untouched
    assert m.has_normals
    return want


def test_c2_sphere_three_passes(ren):
    (d, _), _ = c2_scene()
    want = check(ren, d["pos"], d["idx"])
    np.testing.assert_array_equal(bits(want), bits(d["nrm"]))  # the generator's numpy normals


def test_c1_sphere_two_passes_and_int64_indices(ren):
    from rendertoy_amd import scenes
    d = scenes.icosphere(64, 1000)
    check(ren, d["pos"], d["idx"])
    check(ren, d["pos"], d["idx"], int64=True)


def test_small_meshes_one_pass_repeats_unused_out_of_range(ren):
    rng = np.random.default_rng(12)
    pos = rng.uniform(-2, 2, (200, 3)).astype(f32)
    idx = rng.integers(0, 150, (333, 3)).astype(np.int32)     # 150..199 unused
    idx[:40, 2] = idx[:40, 0]                                  # repeated corners
    idx[::9, 1] = 5000                                         # out of range -> vertex 199
    idx[::11, 0] = -7
    want = check(ren, pos, idx)
    assert not want[150:199].any()


def test_non_indexed_and_degenerate_cases(ren):
    rng = np.random.default_rng(13)
    pos = rng.uniform(-1, 1, (3 * 1000 + 2, 3)).astype(f32)
    check(ren, pos, None)
    # one vertex, one degenerate triangle (0, 0, 0): zero normal
    want = check(ren, np.array([[0.5, -0.25, 2.0]], f32), np.zeros((1, 3), np.int32))
    assert not want.any()
    # collinear triangle among good ones
    p = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [0, 1, 0], [9, 9, 9]], f32)
    check(ren, p, np.array([[0, 1, 2], [0, 1, 3]], np.int32))


def test_four_pass_vertex_ids(ren):
    # 2^24 + 5 vertices: vertex ids need 25 bits
    rng = np.random.default_rng(14)
    nv = (1 << 24) + 5
    pos = rng.uniform(-1, 1, (nv, 3)).astype(f32)
    idx = rng.integers(0, nv, (1 << 20, 3)).astype(np.int32)
    idx[:, 0] = np.arange(1 << 20) * 16                        # ids spread over the full range
    idx[-1] = [nv - 1, nv - 2, nv - 3]
    check(ren, pos, idx)


def test_repeatable_and_feeds_the_render(ren):
    from rendertoy_amd import scenes
    d = scenes.icosphere(48, 321)
    # mesh without normals, given its normals on the device
    m = ren.mesh_from_arrays(d["pos"], d["idx"], None, d["uv"])
    m.compute_normals()
    first = record_floats(m)[:, 4:7].copy()
    m.compute_normals()
    np.testing.assert_array_equal(bits(record_floats(m)[:, 4:7]), bits(first))
    host = ren.mesh_from_arrays(d["pos"], d["idx"], d["nrm"], d["uv"])
    cam = camera()
    light = (np.array([1, 1, 1], f32) / f32(np.sqrt(3.0))).astype(f32)
    a = ren.Raycaster([m]).render(cam, 256, 192, light=light).cpu().numpy()
    b = ren.Raycaster([host]).render(cam, 256, 192, light=light).cpu().numpy()
    np.testing.assert_array_equal(bits(a), bits(b))
    assert (a > 0).any()


def test_bad_arguments_raise(ren):
    from rendertoy_amd import scenes
    d = scenes.icosphere(2, 1)
    m = ren.mesh_from_arrays(d["pos"], d["idx"])
    with pytest.raises(TypeError):
        ren.compute_normals(m, out=torch.empty((3, 3), dtype=torch.float32, device="cuda"))
    with pytest.raises(TypeError):
        ren.compute_normals(m, out=torch.empty((len(d["pos"]), 3), dtype=torch.float64, device="cuda"))


def test_pack_equals_host_packing_and_round_trips(ren):
    from rendertoy_amd import scenes
    d = scenes.icosphere(64, 1000)
    dev = torch.device("cuda")
    for nrm, uv in ((d["nrm"], d["uv"]), (None, d["uv"]), (d["nrm"], None), (None, None)):
        host = ren.mesh_from_arrays(d["pos"], d["idx"], nrm, uv)
        m = ren.mesh_from_tensors(torch.from_numpy(d["pos"]).to(dev), torch.from_numpy(d["idx"]).to(dev),
                                  None if nrm is None else torch.from_numpy(nrm).to(dev),
                                  None if uv is None else torch.from_numpy(uv).to(dev))
        assert m.has_normals == (nrm is not None) and m.has_uvs == (uv is not None)
        np.testing.assert_array_equal(m.vertices.get().view(np.uint8), host.vertices.get().view(np.uint8))
        np.testing.assert_array_equal(m.indices.get(), host.indices.get())
        back = ren.mesh_tensors(m)
        np.testing.assert_array_equal(bits(back["positions"].cpu().numpy()), bits(d["pos"]))
        np.testing.assert_array_equal(bits(back["normals"].cpu().numpy()),
                                      bits(np.zeros_like(d["pos"]) if nrm is None else nrm))
        np.testing.assert_array_equal(bits(back["uvs"].cpu().numpy()),
                                      bits(np.zeros_like(d["uv"]) if uv is None else uv))
        np.testing.assert_array_equal(back["indices"].cpu().numpy(), d["idx"])


def test_tensor_mesh_builds_the_same_scene(ren):
    from rendertoy_amd import scenes
    d = scenes.icosphere(32, 44)
    dev = torch.device("cuda")
    m = ren.mesh_from_tensors(torch.from_numpy(d["pos"]).to(dev),
                              torch.from_numpy(d["idx"].astype(np.int64)).to(dev))
    m.compute_normals()
    host = ren.mesh_from_arrays(d["pos"], d["idx"], d["nrm"], None)
    a, b = ren.Raycaster([m]), ren.Raycaster([host])
    for k in ("_nodes4", "_tris", "_shade"):
        assert torch.equal(getattr(a, k), getattr(b, k)), k
    np.testing.assert_array_equal(bits(ren.mesh_tensors(m)["normals"].cpu().numpy()), bits(d["nrm"]))


def test_pack_rejects_bad_arguments(ren):
    from rendertoy_amd import _native
    lib = _native.load()
    buf = torch.empty(1024, dtype=torch.uint8, device="cuda")
    p = buf.data_ptr()
    s = _native.stream_handle()
    assert lib.rt_mesh_pack(None, None, None, 4, p, s) == -1           # positions missing
    assert lib.rt_mesh_pack(p + 2, None, None, 4, p + 512, s) == -1      # misaligned positions
    assert lib.rt_mesh_pack(p, None, None, 4, p + 520, s) == -1          # records not 16-byte aligned
    assert lib.rt_mesh_unpack(p + 8, 4, p + 512, None, None, s) == -1    # records not 16-byte aligned
    assert lib.rt_mesh_pack(p, None, None, 0, None, s) == 0              # nothing to do
    with pytest.raises(TypeError):
        ren.mesh_from_tensors(torch.zeros((4, 3), dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):
        ren.mesh_from_tensors(torch.zeros((4, 3), device="cuda"), normals=torch.zeros((5, 3), device="cuda"))
