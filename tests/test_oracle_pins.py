"""Pins of two oracle functions the device checks lean on (CPU only):

- `closest64` culled by the hierarchy (tw_closest64_bvh) is the FP64 reference of the C1/C2 GPU
  parity tests and of tests/golden/c1_tile.npz. Its culling must never change a result: on trees
  with many internal levels it must equal the brute-force FP64 closest hit (tw_closest64_brute)
  bit for bit -- t, u, v, prim and every flag -- on random, camera, edge, vertex and tie rays.
- The quad-lane camera emulation (`TwinScene.trace_camera(quad=True)`) supplies the hit counts of
  the roofline's algorithmic bytes and the reference of the large-scene GPU test. Its hits must
  equal the per-ray walk and FP32 brute force, and its per-ray counts (node records and triangles
  a ray's own quad tested) must equal an independent restatement of the 16-ray packet walk
  written here from the kernel's rules (rendertoy_amd/csrc/rt_common.h, traverse_quad_lanes).
"""
import math

import numpy as np
import pytest

import oracle
from helpers import camera, random_rays, twin_from_meshes
from rendertoy_amd import scenes

EMPTY = 0x7FFFFFFF
F32 = np.float32


# ---------------------------------------------------------------------------------------------
# closest64 over the hierarchy == brute force
# ---------------------------------------------------------------------------------------------

def tree_depth(tw):
    """Internal levels of the twin's binary tree (root = level 1)."""
    if tw.root < 0:
        return 0
    refs = tw.nodes[:, 12:14].view(np.int32)
    depth, todo = 0, [(0, 1)]
    while todo:
        x, dpt = todo.pop()
        depth = max(depth, dpt)
        todo += [(int(c), dpt + 1) for c in refs[x] if c >= 0]
    return depth


def pin_scene():
    """An f=24 displaced icosphere, the ground plane, and a copy of 60 of the sphere's triangles
    as a third mesh (exact depth ties between meshes)."""
    s = scenes.icosphere(24, 2400)
    g = scenes.ground_plane()
    dup = {"pos": s["pos"], "idx": s["idx"][:60].copy(), "nrm": s["nrm"], "uv": s["uv"]}
    return [s, g, dup]


def constructed_rays(meshes, rng, eye=(0.3, 0.5, 2.6)):
    """Rays from `eye` through triangle vertices, edge midpoints and centroids of sphere
    triangles (shared edges and vertices: EDGE flags), and through centroids of the duplicated
    triangles (exact depth ties: DEPTH_TIE)."""
    s = meshes[0]
    tri = s["pos"][s["idx"]].astype(np.float64)  # [T, 3, 3]
    pick = rng.choice(len(tri), 400, replace=False)
    pts = [tri[pick, 0], 0.5 * (tri[pick, 0] + tri[pick, 1]), tri[pick].mean(1),
           tri[:60].mean(1)]
    p = np.concatenate(pts)
    o = np.tile(np.asarray(eye, np.float64), (len(p), 1))
    d = p - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(np.float32), d.astype(np.float32)


def test_closest64_bvh_equals_brute_force_on_a_deep_tree():
    meshes = pin_scene()
    tw = twin_from_meshes(meshes)
    assert tw.root == 0 and tree_depth(tw) >= 10, "the pin needs a multi-level hierarchy"
    rng = np.random.default_rng(77)
    o1, d1 = random_rays(20000, 78)
    cam = camera((0.3, 0.5, 2.6), (0, 0, 0), (0, 1, 0), math.pi / 4, 4 / 3)
    o2, d2 = oracle.twin_generate_rays(cam.as_array(), 80, 60)
    o3, d3 = constructed_rays(meshes, rng)
    # rays starting inside the sphere (many boxes contain the origin)
    o4 = rng.uniform(-0.2, 0.2, (3000, 3)).astype(np.float32)
    d4 = rng.standard_normal((3000, 3))
    d4 = (d4 / np.linalg.norm(d4, axis=1, keepdims=True)).astype(np.float32)
    o = np.concatenate([o1, o2, o3, o4])
    d = np.concatenate([d1, d2, d3, d4])
    for tmin in (0.0, 0.5):
        bvh = tw.closest64(o, d, tmin=tmin)
        brute = tw.closest64(o, d, tmin=tmin, brute=True)
        np.testing.assert_array_equal(bvh["prim"], brute["prim"])
        np.testing.assert_array_equal(bvh["t"].view(np.uint64), brute["t"].view(np.uint64))
        np.testing.assert_array_equal(bvh["u"].view(np.uint64), brute["u"].view(np.uint64))
        np.testing.assert_array_equal(bvh["v"].view(np.uint64), brute["v"].view(np.uint64))
        np.testing.assert_array_equal(bvh["edge"], brute["edge"])
        # the constructed rays exercise every flag and both hits and misses
        e = brute["edge"]
        assert (e & oracle.EDGE).any() and (e & oracle.DEPTH_TIE).any()
        assert (brute["prim"] >= 0).sum() > 10000 and (brute["prim"] < 0).sum() > 1000
    # the duplicated triangles tie exactly: the smaller primitive id (the sphere's) wins
    n0 = len(meshes[0]["idx"]) + len(meshes[1]["idx"])
    tie = tw.closest64(o3[-60:], d3[-60:])
    assert (tie["prim"] < n0).all()
    seen = tie["prim"] < 60  # the ray reaches its duplicated triangle (not hidden behind others)
    assert seen.sum() >= 10 and (tie["edge"][seen] & oracle.DEPTH_TIE).all()


# ---------------------------------------------------------------------------------------------
# quad-lane camera emulation
# ---------------------------------------------------------------------------------------------

@pytest.mark.parametrize("view", [((0.2, 0.4, 2.2), 0.9, 37, 23), ((-1.3, 0.2, 1.9), 0.6, 61, 45),
                                  ((0.0, 2.5, 0.8), 1.2, 29, 31)])
def test_quad_emulation_hits_equal_per_ray_and_brute(view):
    eye, fov, W, H = view
    d = scenes.icosphere(24, 31)
    g = scenes.ground_plane()
    tw = twin_from_meshes([d, g])
    up = (0, 0, 1) if eye[1] > 2 else (0, 1, 0)
    cam = camera(eye, (0, 0, 0), up, fov, W / H)
    quad = tw.trace_camera(cam.as_array(), W, H, quad=True)
    o, dd = oracle.twin_generate_rays(cam.as_array(), W, H)
    per_ray = tw.trace_closest(o, dd)
    np.testing.assert_array_equal(quad["prim"], per_ray["prim"])
    for k in ("t", "u", "v"):
        np.testing.assert_array_equal(quad[k].view(np.uint32), per_ray[k].view(np.uint32))
    bt, bp = tw.brute_closest32(o, dd)
    np.testing.assert_array_equal(quad["prim"], bp)
    assert len(np.unique(bp[bp >= 0])) > 50  # many distinct winners, not one big triangle
    # counts per 4x4 packet: every record the packet fetched was tested by at least one of its rays
    # and the packet fetched every record any of its rays tested
    tiles_x = (W + 3) // 4
    col, row = np.meshgrid(np.arange(W), np.arange(H))
    tile = ((row // 4) * tiles_x + col // 4).ravel()
    for lane_key, wave_key in (("lane_nodes", "wave_nodes"), ("lane_tris", "wave_tris")):
        lane = quad[lane_key].astype(np.int64)
        mx = np.zeros(len(quad[wave_key]), np.int64)
        np.maximum.at(mx, tile, lane)
        sm = np.bincount(tile, lane, minlength=len(mx)).astype(np.int64)
        assert (mx <= quad[wave_key]).all() and (quad[wave_key] <= sm).all()


def fma32(a, b, c):
    """fmaf in float32 via float64 (the product of two float32 is exact in float64)."""
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(F32)


def safe_inv32(x):
    x = np.asarray(x, F32)
    tiny = F32(2.0 ** -80)
    return (F32(1.0) / np.where(np.abs(x) < tiny, np.copysign(tiny, x), x).astype(F32)).astype(F32)


def mt32_np(rec, o, d):
    """Two-sided Moller-Trumbore on a (v0|prim, e1, e2) record for rays o, d [R, 3] in FP32 with
    the kernel's operation order: returns (hit mask, t)."""
    a, b, c = rec[0:3], rec[4:7], rec[8:11]
    px = fma32(d[:, 1], c[2], -(d[:, 2] * c[1]).astype(F32))
    py = fma32(d[:, 2], c[0], -(d[:, 0] * c[2]).astype(F32))
    pz = fma32(d[:, 0], c[1], -(d[:, 1] * c[0]).astype(F32))
    det = fma32(b[0], px, fma32(b[1], py, (b[2] * pz).astype(F32)))
    sg = np.where(det > 0, F32(1), F32(-1)).astype(F32)
    ad = (det * sg).astype(F32)
    sx, sy, sz = ((o[:, k] - a[k]).astype(F32) for k in range(3))
    un = fma32(sx, px, fma32(sy, py, (sz * pz).astype(F32)))
    us = (un * sg).astype(F32)
    ok = (det != 0) & (us >= 0) & (us <= ad)
    qx = fma32(sy, b[2], -(sz * b[1]).astype(F32))
    qy = fma32(sz, b[0], -(sx * b[2]).astype(F32))
    qz = fma32(sx, b[1], -(sy * b[0]).astype(F32))
    vn = fma32(d[:, 0], qx, fma32(d[:, 1], qy, (d[:, 2] * qz).astype(F32)))
    vs = (vn * sg).astype(F32)
    ok &= (vs >= 0) & ((us + vs).astype(F32) <= ad)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = (F32(1.0) / det).astype(F32)
        t = (fma32(c[0], qx, fma32(c[1], qy, (c[2] * qz).astype(F32))) * inv).astype(F32)
    return ok, t


def quad_packet_walk(nodes4, tris, root, o, d):
    """The closest-hit walk of one quad-lane packet (at most 16 camera rays, tmin 0, tmax inf):
    the packet visits one node at a time; at a node every ray of its mask tests the 4 child boxes
    (robust slabs, culled by the ray's best t); internal children some ray entered are visited
    nearest first by the entry distance of the lowest ray of the mask (ties: child index), the
    others pushed with their ray masks; leaf children some ray entered are tested in the same step
    by the rays that entered them, in child order; an empty step pops the stack. Returns per ray
    (prim, node records tested, triangles tested) and the packet's (node records, triangle
    records) fetched."""
    R = len(o)
    best = np.full(R, np.inf, F32)
    prim = np.full(R, -1, np.int64)
    nv = np.zeros(R, np.int64)
    nt = np.zeros(R, np.int64)
    fetched = [0, 0]
    inv = [safe_inv32(d[:, k]) for k in range(3)]
    org = [(o[:, k] * inv[k]).astype(F32) for k in range(3)]

    def leaf(ref, mask):
        u = (~np.int64(ref)) & 0xFFFFFFFF
        first, cnt = int(u & 0x1FFFFFFF), int(u >> 29) + 1
        fetched[1] += cnt
        for s in range(first, first + cnt):
            rec = tris[s]
            p = int(rec[3:4].view(np.int32)[0])
            hit, t = mt32_np(rec, o, d)
            nt[mask] += 1
            for r in np.nonzero(mask & hit & (t > 0))[0]:
                if t[r] < best[r] or (t[r] == best[r] and p < prim[r]):
                    best[r], prim[r] = t[r], p

    if root == EMPTY:
        return prim, nv, nt, fetched
    if root < 0:
        leaf(root, np.ones(R, bool))
        return prim, nv, nt, fetched
    stack = []
    cur, m = root, np.ones(R, bool)
    scale = F32(1.0000004)
    while True:
        fetched[0] += 1
        nv[m] += 1
        rec = nodes4[cur]
        refs = rec[24:28].view(np.int32)
        hk, tnk = [], []
        for k in range(4):
            s0 = [fma32(rec[8 * j + k], inv[j], -org[j]) for j in range(3)]
            s1 = [fma32(rec[8 * j + 4 + k], inv[j], -org[j]) for j in range(3)]
            lo = [np.minimum(a, b) for a, b in zip(s0, s1)]
            hi = [np.maximum(a, b) for a, b in zip(s0, s1)]
            tn = np.maximum(np.maximum(lo[0], lo[1]), np.maximum(lo[2], F32(0)))
            tf = (np.minimum(np.minimum(hi[0], hi[1]), hi[2]) * scale).astype(F32)
            hk.append(m & (refs[k] != EMPTY) & (tn <= np.minimum(tf, best)))
            tnk.append(tn)
        rep = int(np.argmax(m))
        key = [float(tnk[k][rep]) if hk[k][rep] else math.inf for k in range(4)]
        live = [k for k in sorted(range(4), key=lambda k: (key[k], k)) if refs[k] >= 0 and hk[k].any()]
        stack += [(int(refs[k]), hk[k]) for k in reversed(live[1:])]
        nxt = (int(refs[live[0]]), hk[live[0]]) if live else (stack.pop() if stack else None)
        for k in range(4):
            if refs[k] < 0 and refs[k] != EMPTY and hk[k].any():
                leaf(int(refs[k]), hk[k])
        if nxt is None:
            return prim, nv, nt, fetched
        cur, m = nxt


@pytest.mark.parametrize("view", [((0.4, 0.7, 2.0), 0.9, 21, 17), ((1.7, -0.1, 1.2), 0.7, 18, 13)])
def test_quad_emulation_counts_equal_independent_walk(view):
    eye, fov, W, H = view
    d = scenes.icosphere(6, 61)
    g = scenes.ground_plane()
    tw = twin_from_meshes([d, g])
    assert tw.root == 0 and tree_depth(tw) >= 6
    cam = camera(eye, (0, 0, 0), (0, 1, 0), fov, W / H)
    quad = tw.trace_camera(cam.as_array(), W, H, quad=True)
    o, dd = oracle.twin_generate_rays(cam.as_array(), W, H)
    tiles_x, tiles_y = (W + 3) // 4, (H + 3) // 4
    for ty in range(tiles_y):
        for tx in range(tiles_x):
            cols = [tx * 4 + (q & 3) for q in range(16)]
            rows = [ty * 4 + (q >> 2) for q in range(16)]
            pix = np.array([r * W + c for c, r in zip(cols, rows) if c < W and r < H])
            prim, nv, nt, fetched = quad_packet_walk(tw.nodes4, tw.tris, tw.root, o[pix], dd[pix])
            np.testing.assert_array_equal(quad["prim"][pix], prim)
            np.testing.assert_array_equal(quad["lane_nodes"][pix], nv)
            np.testing.assert_array_equal(quad["lane_tris"][pix], nt)
            t = ty * tiles_x + tx
            assert (quad["wave_nodes"][t], quad["wave_tris"][t]) == tuple(fetched)


def test_identical_rays_walk_as_one():
    """A packet of 16 bit-identical rays (fov so small that every direction rounds to the view
    axis) visits exactly the records of the same ray alone."""
    d = scenes.icosphere(12, 71)
    tw = twin_from_meshes([d, scenes.ground_plane()])
    cam = camera((0.3, 0.2, 2.5), (0.01, 0.02, 0.0), (0, 1, 0), 1e-9, 1.0)
    many = tw.trace_camera(cam.as_array(), 4, 4, quad=True)
    one = tw.trace_camera(cam.as_array(), 1, 1, quad=True)
    o, dd = oracle.twin_generate_rays(cam.as_array(), 4, 4)
    assert (dd == dd[0]).all()
    for k in ("prim", "lane_nodes", "lane_tris"):
        assert (many[k] == one[k][0]).all()
    assert many["wave_nodes"][0] == one["wave_nodes"][0] == one["lane_nodes"][0]
    assert many["wave_tris"][0] == one["wave_tris"][0] == one["lane_tris"][0]
    assert one["prim"][0] >= 0 and one["lane_nodes"][0] > 3
