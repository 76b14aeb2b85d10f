"""Pins of the path-sampling oracle (CPU): Philox4x32-10 against the published known-answer
vectors of the Random123 distribution and an independent numpy restatement; the polynomial
cos/sin of 2 pi x against numpy; cosine-weighted directions against the moments of the cosine
distribution; and the twin's path pipeline against scenes with a closed-form answer (a flat quad:
every path escapes after its first bounce; the inside of a closed sphere: no path escapes)."""
import math

import numpy as np
import pytest

import oracle
from helpers import twin_from_meshes, camera

KATS = [  # Random123 kat_vectors, philox4x32 with 10 rounds: counter, key, output
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,want", KATS)
def test_philox_known_answers(ctr, key, want):
    assert tuple(int(x) for x in oracle.twin_philox(np.array(ctr), key)[0]) == want
    assert tuple(int(x) for x in oracle.philox_np(np.array(ctr), key)[0]) == want


def test_philox_twin_equals_numpy_restatement():
    rng = np.random.default_rng(5)
    ctr = rng.integers(0, 2 ** 32, (5000, 4), dtype=np.uint64).astype(np.uint32)
    for key in ((0x5EED, 0), (0x5EED, 7), (0xDEADBEEF, 0x12345678)):
        np.testing.assert_array_equal(oracle.twin_philox(ctr, key), oracle.philox_np(ctr, key))


def test_cos_sin_2pi_polynomial():
    x = (np.arange(1 << 16, dtype=np.float64) / (1 << 16)).astype(np.float32)
    c, s = oracle.twin_cos_sin_2pi(x)
    ang = 2 * np.pi * x.astype(np.float64)
    assert np.abs(c - np.cos(ang)).max() < 3e-7
    assert np.abs(s - np.sin(ang)).max() < 3e-7
    # exact quadrant starts
    q = np.array([0.0, 0.25, 0.5, 0.75], np.float32)
    cq, sq = oracle.twin_cos_sin_2pi(q)
    np.testing.assert_array_equal(cq, [1, 0, -1, 0])
    np.testing.assert_array_equal(sq, [0, 1, 0, -1])


def test_cosine_directions_follow_the_cosine_law():
    rng = np.random.default_rng(6)
    n = 200000
    nrm = rng.standard_normal((n, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    u0 = rng.random(n).astype(np.float32)
    u1 = rng.random(n).astype(np.float32)
    d = oracle.twin_cosine_dir(nrm, u0, u1).astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=3e-7)
    cos = np.einsum("ij,ij->i", d, nrm.astype(np.float64))
    assert cos.min() > -1e-6
    # cos(theta) = sqrt(1 - u0) exactly in the basis: E[cos] = 2/3, E[cos^2] = 1/2
    np.testing.assert_allclose(cos, np.sqrt(1 - u0.astype(np.float64)), atol=2e-6)
    assert abs(cos.mean() - 2 / 3) < 3e-3 and abs((cos ** 2).mean() - 0.5) < 3e-3
    # the azimuth is uniform: the tangential part has no preferred direction
    t = d - cos[:, None] * nrm
    assert np.abs(t.mean(0)).max() < 5e-3


def quad():
    return {"pos": np.array([(-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0)], np.float32),
            "idx": np.array([(0, 1, 2), (0, 2, 3)], np.int32), "nrm": None, "uv": None}


def test_paths_flat_quad_escapes_after_one_bounce():
    tw = twin_from_meshes([quad()])
    W, H = 48, 40
    cam = camera((0.2, 0.3, 2.5), (0, 0, 0), (0, 1, 0), math.pi / 3, W / H)
    rad, smp, cnt = tw.paths(cam.as_array(), W, H, spp=4, bounces=2, albedo=0.7, sky=1.0)
    hit = smp == np.float32(0.7)
    assert set(np.unique(smp).tolist()) == {np.float32(0.7).item(), 1.0}
    assert cnt[0] == W * H * 4 and cnt[1] == hit.sum() and cnt[2] == 0
    np.testing.assert_array_equal(rad, smp.reshape(-1, 4).sum(1, dtype=np.float32) / np.float32(4))
    assert 0.2 < hit.mean() < 0.9


def test_paths_inside_a_closed_sphere_never_escape():
    from rendertoy_amd import scenes
    unit, idx = scenes.geodesic_sphere(6)
    sph = {"pos": (unit * 2.0).astype(np.float32), "idx": idx, "nrm": None, "uv": None}
    tw = twin_from_meshes([sph])
    W, H = 32, 24
    cam = camera((0.1, 0.2, 0.3), (0, 0, -1), (0, 1, 0), math.pi / 3, W / H)
    for bounces in (0, 1, 3):
        rad, smp, cnt = tw.paths(cam.as_array(), W, H, spp=2, bounces=bounces)
        assert (smp == 0).all() and (rad == 0).all()
        assert list(cnt) == [W * H * 2] * (bounces + 1)   # every ray hits the enclosure


def test_paths_are_seeded_and_sample_counts_are_exact():
    from rendertoy_amd import scenes
    a = scenes.icosphere(8, 71)
    tw = twin_from_meshes([a, quad()])
    W, H = 40, 32
    cam = camera((0.3, 0.8, 2.4), (0, 0, 0), (0, 1, 0), math.pi / 4, W / H)
    r1 = tw.paths(cam.as_array(), W, H, spp=3, bounces=2, frame=1)
    r2 = tw.paths(cam.as_array(), W, H, spp=3, bounces=2, frame=1)
    r3 = tw.paths(cam.as_array(), W, H, spp=3, bounces=2, frame=2)
    np.testing.assert_array_equal(r1[1], r2[1])
    assert not np.array_equal(r1[1], r3[1])
    vals = {0.0, 1.0, np.float32(0.7).item(), (np.float32(0.7) * np.float32(0.7)).item()}
    assert set(np.unique(r1[1]).tolist()) <= vals
    assert r1[2][0] == W * H * 3 and r1[2][1] >= r1[2][2] > 0
