"""Pins of the oracle's per-vertex normals (tw_vertex_normals; Mesh.compute_normals, whose reference
body rendering/_modeling.py:48-49 is a stub). CPU only.

Definition: N(v) = n / |n|, n = FP64 sum of cross(P1 - P0, P2 - P0) over the triangles using v
(corner-0 pairs first, then corner 1, then corner 2, ascending triangle), rounded once to float32;
unused vertices get 0. Pinned by:
- the scene generator's independent numpy statement (np.cross + np.add.at), bit for bit;
- geometry: a flat fan gives exactly +z (and -z reversed), the normal at a corner shared by two
  perpendicular faces leans by the ratio of their areas (area weighting, not a plain average), the
  vertices of a regular geodesic sphere point radially (exactly at its 5-fold symmetric vertices);
- exact invariances: winding reversal negates every bit, scaling by 2 changes nothing, an index
  list 0..3T-1 equals the non-indexed mesh, out-of-range indices act as n_vertices - 1."""
import math

import numpy as np
import pytest

from rendertoy_amd import scenes

f32 = np.float32


def normals(pos, idx=None):
    import oracle
    return oracle.vertex_normals(pos, idx)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.mark.parametrize("f,seed", [(2, 1), (8, 301), (32, 7)])
def test_twin_equals_generator_numpy(f, seed):
    d = scenes.icosphere(f, seed)
    np.testing.assert_array_equal(bits(normals(d["pos"], d["idx"])), bits(scenes.vertex_normals(d["pos"], d["idx"])))


def test_twin_equals_numpy_on_random_mesh_with_repeats_and_unused():
    rng = np.random.default_rng(11)
    pos = rng.uniform(-3, 3, (500, 3)).astype(f32)
    idx = rng.integers(0, 400, (900, 3)).astype(np.int32)   # vertices 400..499 unused
    idx[:50, 1] = idx[:50, 0]                               # repeated corners (degenerate faces)
    n = normals(pos, idx)
    np.testing.assert_array_equal(bits(n), bits(scenes.vertex_normals(pos, idx)))
    assert not n[400:].any()


def test_flat_fan_is_exactly_z():
    # a CCW fan around the origin in the plane z = 0.25
    k = 9
    ang = np.arange(k) * (2 * math.pi / k)
    pos = np.concatenate([[[0, 0, 0.25]], np.stack([np.cos(ang), np.sin(ang), np.full(k, 0.25)], 1),
                          [[5, 5, 5]]]).astype(f32)                     # last vertex unused
    idx = np.array([[0, 1 + i, 1 + (i + 1) % k] for i in range(k)], np.int32)
    n = normals(pos, idx)
    np.testing.assert_array_equal(n[:-1], np.tile(f32([0, 0, 1]), (k + 1, 1)))
    np.testing.assert_array_equal(n[-1], f32([0, 0, 0]))
    np.testing.assert_array_equal(normals(pos, idx[:, ::-1].copy())[:-1], np.tile(f32([0, 0, -1]), (k + 1, 1)))


@pytest.mark.parametrize("a,b,c,e", [(2.0, 1.0, 1.0, 1.0), (1.0, 1.0, 3.0, 0.5), (4.0, 0.25, 1.0, 1.0)])
def test_area_weighting_at_a_shared_corner(a, b, c, e):
    # face 1 in z = 0 (normal +z, area ab/2), face 2 in x = 0 (normal +x, area ce/2), sharing vertex 0
    pos = np.array([[0, 0, 0], [a, 0, 0], [0, b, 0], [0, c, 0], [0, 0, e]], f32)
    idx = np.array([[0, 1, 2], [0, 3, 4]], np.int32)
    n = normals(pos, idx)[0].astype(np.float64)
    # the normal at vertex 0 lies in the xz-plane at atan2(area1, area2) from +x
    assert n[1] == 0.0
    assert abs(math.atan2(n[2], n[0]) - math.atan2(a * b, c * e)) < 2e-7
    assert abs(np.linalg.norm(n) - 1.0) < 1e-7


@pytest.mark.parametrize("f", [4, 16, 64])
def test_regular_sphere_normals_are_radial(f):
    unit, idx = scenes.geodesic_sphere(f)
    pos = unit.astype(f32)
    n = normals(pos, idx).astype(np.float64)
    r = pos.astype(np.float64)
    r /= np.linalg.norm(r, axis=1, keepdims=True)
    ang = np.arctan2(np.linalg.norm(np.cross(n, r), axis=1), (n * r).sum(1))  # well-conditioned near 0
    assert ang.max() < 0.3 / f
    valence = np.bincount(idx.ravel(), minlength=len(pos))
    assert (valence == 5).sum() == 12
    assert ang[valence == 5].max() < 1e-6


def test_winding_reversal_negates_bits():
    d = scenes.icosphere(8, 5)
    a = normals(d["pos"], d["idx"])
    b = normals(d["pos"], d["idx"][:, [0, 2, 1]].copy())
    np.testing.assert_array_equal(bits(-a), bits(b))


def test_scaling_by_two_is_exact():
    d = scenes.icosphere(8, 6)
    np.testing.assert_array_equal(bits(normals(d["pos"], d["idx"])),
                                  bits(normals((d["pos"] * f32(2)).astype(f32), d["idx"])))


def test_non_indexed_equals_identity_indices_and_tail_is_zero():
    rng = np.random.default_rng(3)
    pos = rng.uniform(-1, 1, (3 * 40 + 2, 3)).astype(f32)          # 2 trailing vertices unused
    a = normals(pos, None)
    b = normals(pos, np.arange(120, dtype=np.int32).reshape(-1, 3))
    np.testing.assert_array_equal(bits(a), bits(b))
    assert not a[120:].any()
    # each vertex of a non-indexed mesh carries its own face's unit normal
    p = pos[:120].reshape(-1, 3, 3).astype(np.float64)
    fn = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    np.testing.assert_allclose(a[:120].reshape(-1, 3, 3), np.repeat(fn[:, None], 3, 1), atol=1e-6)


def test_out_of_range_indices_clamp_to_last_vertex():
    rng = np.random.default_rng(4)
    pos = rng.uniform(-1, 1, (30, 3)).astype(f32)
    idx = rng.integers(0, 30, (40, 3)).astype(np.int32)
    bad = idx.copy()
    bad[::7, 0] = 1000
    bad[::5, 2] = -3
    fixed = np.where((bad < 0) | (bad >= 30), 29, bad).astype(np.int32)
    np.testing.assert_array_equal(bits(normals(pos, bad)), bits(normals(pos, fixed)))


def test_compute_normals_needs_the_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("HIP device present")
    import rendertoy_amd as ren
    with pytest.raises(RuntimeError):
        ren.compute_normals(ren.Mesh.__new__(ren.Mesh))
