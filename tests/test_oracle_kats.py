"""Pins of the oracle's building blocks against closed forms and textbook definitions (CPU only).

The reference ray caster is a stub with no tests or golden data (rendering/_raycaster.py:30-36),
so every expected value here comes from a definition: bit interleaving, Moller-Trumbore on
hand-checkable geometry and Cramer's rule in FP64, the look_at/perspective camera of
rendering/_core.py:528-548 with the viewport of rendering/_raster.py:126-129, and the OpenCL 1.2
linear filter / the reference's point-sampling rule (rendering/_core.py:94-96)."""
import math

import numpy as np
import pytest

import oracle
from helpers import twin_from_meshes, camera


# ---------------------------------------------------------------------------------- Morton

def test_expand10_known_values():
    L = oracle.lib()
    assert L.tw_expand10(0x3FF) == 0x09249249
    assert L.tw_expand10(1) == 1
    assert L.tw_expand10(2) == 8
    assert L.tw_expand10(0x200) == 1 << 27
    assert L.tw_expand10(0x155) == 0x01041041  # bits 0,2,4,6,8 -> 0,6,12,18,24


def test_morton3_known_values():
    L = oracle.lib()
    assert L.tw_morton3(1, 0, 0) == 4
    assert L.tw_morton3(0, 1, 0) == 2
    assert L.tw_morton3(0, 0, 1) == 1
    assert L.tw_morton3(1023, 1023, 1023) == 0x3FFFFFFF
    assert L.tw_morton3(1023, 0, 0) == 0x24924924


def test_expand10_matches_bit_definition_for_all_inputs():
    L = oracle.lib()
    ref = oracle.expand10_np(np.arange(1024))
    got = np.array([L.tw_expand10(i) for i in range(1024)], np.uint32)
    np.testing.assert_array_equal(got, ref)


def test_morton_order_is_bitwise_interleave_order():
    rng = np.random.default_rng(1)
    q = rng.integers(0, 1024, (5000, 3))
    codes = oracle.morton_np(q[:, 0], q[:, 1], q[:, 2])
    # build the interleaved bit string explicitly (x2 y2 z2 ... x0 y0 z0 from the top)
    s = ["".join(f"{(x >> b) & 1}{(y >> b) & 1}{(z >> b) & 1}" for b in range(9, -1, -1)) for x, y, z in q]
    np.testing.assert_array_equal(codes, np.array([int(b, 2) for b in s], np.uint32))


# ---------------------------------------------------------------------------------- Moller-Trumbore

def one_triangle(v0, v1, v2):
    pos = np.array([v0, v1, v2], np.float32)
    return twin_from_meshes([{"pos": pos, "idx": np.array([[0, 1, 2]], np.int32), "nrm": None, "uv": None}])


def test_mt_axis_aligned_triangle_closed_form():
    tw = one_triangle((-1, -1, 0), (1, -1, 0), (0, 1, 0))
    # P = v0 + u e1 + v e2 with e1 = (2,0,0), e2 = (1,2,0): v = (y+1)/2, u = (x+1-v)/2
    xs = np.array([0.0, 0.25, -0.3, 0.1], np.float32)
    ys = np.array([0.0, -0.5, -0.2, 0.5], np.float32)
    o = np.stack([xs, ys, np.full(4, 5, np.float32)], 1)
    d = np.tile(np.array([0, 0, -1], np.float32), (4, 1))
    h = tw.trace_closest(o, d)
    v = (ys.astype(np.float64) + 1) / 2
    u = (xs.astype(np.float64) + 1 - v) / 2
    assert (h["prim"] == 0).all()
    np.testing.assert_allclose(h["t"], 5.0, rtol=0, atol=1e-6)
    np.testing.assert_allclose(h["u"], u, atol=1e-6)
    np.testing.assert_allclose(h["v"], v, atol=1e-6)


def test_mt_two_sided_parallel_vertex_and_interval():
    tw = one_triangle((-1, -1, 0), (1, -1, 0), (0, 1, 0))
    o = np.array([[0, 0, -5], [-2, 0.5, 0], [0, 1, 3], [-1, -1, 3], [0, 0, 5], [0, 0, 5]], np.float32)
    d = np.array([[0, 0, 1], [1, 0, 0], [0, 0, -1], [0, 0, -1], [0, 0, -1], [0, 0, -1]], np.float32)
    h = tw.trace_closest(o, d)
    assert h["prim"][0] == 0 and abs(h["t"][0] - 5) < 1e-6        # back face: two-sided
    assert h["prim"][1] == -1                                      # parallel, in-plane: det = 0
    assert h["prim"][2] == 0 and h["prim"][3] == 0                 # through vertices v2 and v0
    # interval (tmin, tmax): exclusive at both ends
    h2 = tw.trace_closest(o[4:5], d[4:5], tmin=5.0)
    assert h2["prim"][0] == -1
    h3 = tw.trace_closest(o[4:5], d[4:5], tmax=5.0)
    assert h3["prim"][0] == -1
    h4 = tw.trace_closest(o[4:5], d[4:5], tmin=4.9, tmax=5.1)
    assert h4["prim"][0] == 0


def cramer64(v0, v1, v2, o, d):
    """Ray-triangle intersection by Cramer's rule on o + t d = v0 + u e1 + v e2, FP64."""
    e1, e2 = v1 - v0, v2 - v0
    A = np.stack([-d, e1, e2], 1)
    t, u, v = np.linalg.solve(A, o - v0)
    return t, u, v


def test_mt_matches_cramer_fp64_on_random_triangles():
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(300):
        P = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
        tw = one_triangle(*P)
        target = (P.astype(np.float64).T @ rng.dirichlet([1, 1, 1])).astype(np.float32)
        o = rng.uniform(-3, 3, 3).astype(np.float32)
        d = (target - o)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        h = tw.trace_closest(o[None], d[None])
        t, u, v = cramer64(*P.astype(np.float64), o.astype(np.float64), d.astype(np.float64))
        inside = min(u, v, 1 - u - v) > 1e-3
        if not inside:
            continue
        n = np.cross(P[1].astype(np.float64) - P[0], P[2].astype(np.float64) - P[0])
        if abs(n @ d) / np.linalg.norm(n) < 0.05:
            continue  # grazing: FP32 conditioning, covered by the FP64 error model
        checked += 1
        assert h["prim"][0] == 0
        assert abs(h["t"][0] - t) <= 1e-5 * abs(t)  # refined hit
        assert abs(h["u"][0] - u) < 1e-4 and abs(h["v"][0] - v) < 1e-4
    assert checked > 150


def test_fp64_oracle_closed_form_and_flags():
    P = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32)
    o = np.array([[0.1, 0.2, 4.0], [0.0, -1.0, 4.0], [5.0, 5.0, 4.0]], np.float32)
    d = np.tile(np.array([0, 0, -1], np.float32), (3, 1))
    r = oracle.closest64(P.reshape(1, 9), o, d)
    assert list(r["prim"]) == [0, 0, -1]
    np.testing.assert_allclose(r["t"][:2], 4.0, rtol=1e-15)
    assert r["edge"][0] == 0                   # interior
    assert r["edge"][1] & oracle.EDGE          # on the edge v = 0
    # two coplanar triangles 1e-7 apart in depth: a depth tie for FP32
    Q = np.array([[-1, -1, 0, 1, -1, 0, 0, 1, 0], [-1, -1, 1e-7, 1, -1, 1e-7, 0, 1, 1e-7]], np.float32)
    r2 = oracle.closest64(Q, o[:1], d[:1])
    assert r2["prim"][0] == 1 and r2["edge"][0] & oracle.DEPTH_TIE


# ---------------------------------------------------------------------------------- camera

def test_camera_rays_match_fp64_closed_form():
    W, H, fov = 64, 48, math.pi / 3
    eye, target, up = (0.3, 0.2, 3.0), (0.0, 0.1, 0.0), (0.0, 1.0, 0.0)
    cam = camera(eye, target, up, fov, W / H)
    o, d = oracle.twin_generate_rays(cam.as_array(), W, H)
    ref = oracle.camera_dirs_fp64(eye, target, up, fov, W, H)
    np.testing.assert_allclose(d, ref, atol=3e-7)
    np.testing.assert_array_equal(o, np.tile(np.array(eye, np.float32), (W * H, 1)))


def test_camera_invariants():
    fov = math.pi / 4
    W = H = 101  # odd: a pixel centre on the optical axis
    cam = camera((0, 0, 3), (0, 0, 0), (0, 1, 0), fov, 1.0)
    o, d = oracle.twin_generate_rays(cam.as_array(), W, H)
    centre = d[(H // 2) * W + W // 2]
    np.testing.assert_allclose(centre, [0, 0, -1], atol=1e-7)
    # the top-centre ray's angle to the axis: atan((1 - 1/H) * tan(fov/2))
    top = d[W // 2]
    ang = math.acos(float(-top[2]))
    assert abs(ang - math.atan((1 - 1 / H) * math.tan(fov / 2))) < 1e-6
    assert top[1] > 0                            # row 0 is the top of the image (y flip)
    # left-handed look_at: screen right is world -x for an eye on +z looking at the origin
    right = d[(H // 2) * W + W - 1]
    assert right[0] < 0


# ---------------------------------------------------------------------------------- textures

def test_bilinear_constant_centres_and_wrap():
    rng = np.random.default_rng(5)
    tex = rng.integers(0, 256, (16, 32, 4), dtype=np.uint8)
    # texel centres return the texel exactly
    j, i = np.meshgrid(np.arange(16), np.arange(32), indexing="ij")
    u = ((i + 0.5) / 32).ravel().astype(np.float32)
    v = ((j + 0.5) / 16).ravel().astype(np.float32)
    got = oracle.twin_tex_sample(tex, u, v, True)
    np.testing.assert_allclose(got, tex[j.ravel(), i.ravel(), :3] / 255.0, atol=1e-6)
    # repeat addressing: shifting by whole periods changes nothing
    uu = rng.uniform(-3, 3, 2000).astype(np.float32)
    vv = rng.uniform(-3, 3, 2000).astype(np.float32)
    a = oracle.twin_tex_sample(tex, uu, vv, True)
    ref = oracle.bilinear_fp64(tex, uu, vv)
    np.testing.assert_allclose(a, ref, atol=2e-5)
    # constant texture -> constant
    c = np.full((8, 8, 4), 77, np.uint8)
    np.testing.assert_allclose(oracle.twin_tex_sample(c, uu, vv, True), 77 / 255.0, atol=1e-6)


def test_point_sampling_is_the_reference_rule():
    # rendering/_core.py:94-96: texel[(int)(wrap(v)*H)][(int)(wrap(u)*W)], wrap(c) = frac(c)
    rng = np.random.default_rng(6)
    tex = rng.integers(0, 256, (8, 16, 4), dtype=np.uint8)
    u = rng.uniform(-2, 2, 3000).astype(np.float32)
    v = rng.uniform(-2, 2, 3000).astype(np.float32)
    got = oracle.twin_tex_sample(tex, u, v, False)
    fu = (u - np.floor(u)).astype(np.float32)
    fv = (v - np.floor(v)).astype(np.float32)
    ii = np.minimum((fu * np.float32(16)).astype(np.int64), 15)
    jj = np.minimum((fv * np.float32(8)).astype(np.int64), 7)
    np.testing.assert_allclose(got, tex[jj, ii, :3] / 255.0, atol=1e-7)
