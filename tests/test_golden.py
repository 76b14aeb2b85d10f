"""The committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle reproduces them on the CPU, and the device path matches them on the GPU."""
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LIGHT = (np.array([1.0, 1.0, 1.0], np.float32) / np.float32(np.sqrt(3.0))).astype(np.float32)


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name)))


def u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


def cases():
    import sys
    sys.path.insert(0, GOLDEN)
    import make_golden
    return make_golden


@pytest.mark.parametrize("builder", ["lbvh", "ploc"])
def test_oracle_reproduces_build_fixture(builder):
    g = load(f"{builder}_980.npz")
    c = cases().build_case(builder)
    assert c["geometry"] == str(g["geometry"]), "scene generator changed"
    for k in ("keys", "order", "nodes", "nodes4", "tris", "shade", "bounds", "root"):
        np.testing.assert_array_equal(u32(np.asarray(c[k])), u32(g[k]), err_msg=k)


def test_oracle_reproduces_c1_tile_fixture():
    g = load("c1_tile.npz")
    c = cases().c1_tile_case()
    assert c["geometry"] == str(g["geometry"])
    np.testing.assert_array_equal(c["prim64"], g["prim64"])
    np.testing.assert_array_equal(c["flags"], g["flags"])
    np.testing.assert_allclose(c["t64"], g["t64"], rtol=1e-12)
    np.testing.assert_array_equal(c["prim32"], g["prim32"])
    np.testing.assert_array_equal(u32(c["t32"]), u32(g["t32"]))
    # the fixture itself meets the parity bar: IDs equal except flagged rays, t within 1e-5
    agree = g["prim32"] == g["prim64"]
    assert ((~agree) & (g["flags"] == 0)).sum() == 0
    hit = agree & (g["prim64"] >= 0)
    assert (np.abs(g["t32"][hit] - g["t64"][hit]) <= 1e-5 * g["t64"][hit]).all()


def test_oracle_reproduces_shading_fixture():
    g = load("shade_c2s.npz")
    c = cases().shade_case()
    assert c["geometry"] == str(g["geometry"])
    np.testing.assert_array_equal(u32(c["rgb32"]), u32(g["rgb32"]))
    np.testing.assert_allclose(c["rgb64"], g["rgb64"], rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(c["flags"], g["flags"])
    err = np.abs(g["rgb32"] - g["rgb64"]).max(1)
    assert (err[g["flags"] == 0] <= 1 / 255).all()


@pytest.mark.gpu
@pytest.mark.parametrize("builder", ["lbvh", "ploc"])
def test_device_matches_build_fixture(ren, builder):
    from rendertoy_amd import scenes
    from helpers import gpu_from_meshes, gpu_arrays
    g = load(f"{builder}_980.npz")
    with ren.tuning(builder=builder):
        rc = gpu_from_meshes([scenes.icosphere(7, 4242)])
    a = gpu_arrays(rc)
    assert rc.root == int(g["root"])
    for k in ("keys", "order", "nodes", "nodes4", "tris", "shade", "bounds"):
        np.testing.assert_array_equal(u32(a[k]), u32(g[k]), err_msg=k)


@pytest.mark.gpu
def test_device_matches_c1_tile_fixture(ren):
    from conftest import c1_scene
    from helpers import gpu_from_meshes, camera
    g = load("c1_tile.npz")
    rc = gpu_from_meshes(c1_scene())
    hits = rc.cast_camera(camera(), 512, 512)
    prim = hits["prim"].cpu().numpy()[g["pixels"]]
    t = hits["t"].cpu().numpy()[g["pixels"]]
    np.testing.assert_array_equal(prim, g["prim32"])
    np.testing.assert_array_equal(u32(t), u32(g["t32"]))


@pytest.mark.gpu
def test_device_matches_shading_fixture(ren):
    import torch
    from rendertoy_amd import scenes
    from helpers import gpu_from_meshes, camera
    g = load("shade_c2s.npz")
    tex = scenes.value_noise_rgba8(64, 78)
    rc = gpu_from_meshes([scenes.icosphere(32, 77), scenes.ground_plane()])
    W, H = int(g["W"]), int(g["H"])
    cam = camera((0.4, 0.5, 2.5), (0, 0, 0), (0, 1, 0), math.pi / 4, W / H)
    img = rc.render(cam, W, H, light=LIGHT, textures=[torch.from_numpy(tex).cuda()],
                    mesh_textures=[0, 0]).cpu().numpy().reshape(-1, 3)
    np.testing.assert_array_equal(u32(img), u32(g["rgb32"]))
    err = np.abs(img - g["rgb64"]).max(1)
    assert (err[g["flags"] == 0] <= 1 / 255).all()
