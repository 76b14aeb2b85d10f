"""GPU parity: the HIP kernels of librtx_hip.so against the CPU twin (bit-exact) and the FP64
oracle (IDs exact except counted edge rays, t within 1e-5 relative, RGB within 1/255)."""
import math

import numpy as np
import pytest
import torch

from conftest import c1_scene, c2_scene
from helpers import twin_from_meshes, gpu_from_meshes, gpu_arrays, camera, random_rays, soa_tensors

pytestmark = pytest.mark.gpu

LIGHT = (np.array([1.0, 1.0, 1.0], np.float32) / np.float32(np.sqrt(3.0))).astype(np.float32)


def _u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


def fp64_report(name, prim32, t32, r64):
    """Asserts the FP64 parity bar and prints the counts: IDs equal except on flagged rays (edge,
    depth tie, near miss), t within 1e-5 relative where the IDs agree."""
    import oracle
    mism = prim32 != r64["prim"]
    flag = r64["edge"]
    bad = mism & (flag == 0)
    agree = (prim32 >= 0) & ~mism
    rel = np.abs(t32[agree].astype(np.float64) - r64["t"][agree]) / r64["t"][agree]
    print(f"{name}: rays={len(prim32)} hits={int((prim32 >= 0).sum())} "
          f"id_mismatch={int(mism.sum())} unflagged_mismatch={int(bad.sum())} "
          f"flag_edge={int(((flag & oracle.EDGE) != 0).sum())} "
          f"flag_depth_tie={int(((flag & oracle.DEPTH_TIE) != 0).sum())} "
          f"flag_near_miss={int(((flag & oracle.NEAR_MISS) != 0).sum())} "
          f"t_rel_max={rel.max() if rel.size else 0:.3e} t_rel_mean={rel.mean() if rel.size else 0:.3e}")
    assert int(bad.sum()) == 0, "ID mismatches vs FP64 on unflagged rays"
    assert rel.size == 0 or rel.max() <= 1e-5


def assert_build_equal(rc, tw):
    g = gpu_arrays(rc)
    n = tw.n
    assert rc.root == tw.root
    np.testing.assert_array_equal(g["keys"].view(np.uint32), tw.keys)
    np.testing.assert_array_equal(g["order"].view(np.uint32), tw.order)
    np.testing.assert_array_equal(_u32(g["tris"]), _u32(tw.tris))
    np.testing.assert_array_equal(_u32(g["shade"]), _u32(tw.shade))
    np.testing.assert_array_equal(_u32(g["nodes"]), _u32(tw.nodes))
    np.testing.assert_array_equal(_u32(g["nodes4"]), _u32(tw.nodes4))
    np.testing.assert_array_equal(_u32(g["bounds"]), _u32(tw.bounds))
    assert g["mesh_offsets"][-1] == n


def gpu_trace(rc, o, d, tmin=0.0, tmax=math.inf):
    out = rc.trace_soa(soa_tensors(o, d), tmin=tmin, tmax=tmax)
    return {k: v.cpu().numpy() for k, v in out.items()}


def assert_trace_equal(g, t):
    np.testing.assert_array_equal(g["prim"], t["prim"])
    np.testing.assert_array_equal(_u32(g["t"]), _u32(np.where(t["prim"] >= 0, t["t"], np.float32(np.inf))))
    np.testing.assert_array_equal(_u32(g["u"]), _u32(np.where(t["prim"] >= 0, t["u"], 0)))
    np.testing.assert_array_equal(_u32(g["v"]), _u32(np.where(t["prim"] >= 0, t["v"], 0)))


@pytest.fixture(scope="module")
def c1(ren):
    meshes = c1_scene()
    tw = twin_from_meshes(meshes)
    rc = gpu_from_meshes(meshes)
    torch.cuda.synchronize()
    return meshes, tw, rc


def test_build_c1_matches_twin_bytes(c1):
    meshes, tw, rc = c1
    assert_build_equal(rc, tw)


def test_camera_rays_match_twin(ren, twin):
    cam = camera(aspect=640 / 480)
    rays = ren.Raycaster.generate_rays(cam, 640, 480)
    o, d = twin.twin_generate_rays(cam.as_array(), 640, 480)
    for i, k in enumerate(("dx", "dy", "dz")):
        np.testing.assert_array_equal(_u32(rays[k].cpu().numpy()), _u32(d[:, i]))
    for i, k in enumerate(("ox", "oy", "oz")):
        np.testing.assert_array_equal(rays[k].cpu().numpy(), o[:, i])


def test_trace_c1_primary_matches_twin_and_fp64(c1, twin):
    meshes, tw, rc = c1
    cam = camera()
    o, d = twin.twin_generate_rays(cam.as_array(), 512, 512)
    g = gpu_trace(rc, o, d)
    t = tw.trace_closest(o, d)
    assert_trace_equal(g, t)
    # the camera kernel (wave-coherent traversal) matches its twin emulation and the per-ray twin
    gc = {k: v.cpu().numpy() for k, v in rc.cast_camera(cam, 512, 512).items()}
    tc = tw.trace_camera(cam.as_array(), 512, 512)
    assert_trace_equal(gc, tc)
    assert_trace_equal(gc, t)
    # FP64: identical IDs except flagged rays, t within 1e-5 relative
    fp64_report("C1 primary 512x512", g["prim"], g["t"], tw.closest64(o, d))


def test_trace_random_rays_match_twin(c1):
    meshes, tw, rc = c1
    o, d = random_rays(200_000, 7)
    assert_trace_equal(gpu_trace(rc, o, d), tw.trace_closest(o, d))
    # interval limits
    assert_trace_equal(gpu_trace(rc, o, d, 2.5, 3.2), tw.trace_closest(o, d, 2.5, 3.2))


def test_any_hit_matches_twin(c1):
    meshes, tw, rc = c1
    o, d = random_rays(100_000, 8)
    occ = rc.occluded(soa_tensors(o, d)).cpu().numpy()
    np.testing.assert_array_equal(occ, tw.trace_any(o, d))
    occ = rc.occluded(soa_tensors(o, d), tmin=0.0, tmax=2.7).cpu().numpy()
    np.testing.assert_array_equal(occ, tw.trace_any(o, d, 0.0, 2.7))


def test_ray_cast_aos_hits(c1, ren):
    meshes, tw, rc = c1
    o, d = random_rays(5000, 9)
    rays = np.zeros(len(o), dtype=ren.Ray)
    rf = rays.view(np.float32).reshape(-1, 8)
    rf[:, 0:3] = o
    rf[:, 4:7] = d
    hits = rc.ray_cast(ren.create_buffer_from(rays)).get()
    t = tw.trace_closest(o, d)
    np.testing.assert_array_equal(hits["index"], t["prim"])
    np.testing.assert_array_equal(hits["mesh"], np.where(t["prim"] >= 0, 0, -1))
    np.testing.assert_array_equal(_u32(hits["t"]), _u32(np.where(t["prim"] >= 0, t["t"], np.float32(np.inf))))


def test_render_c1_matches_twin(c1, twin):
    meshes, tw, rc = c1
    cam = camera()
    stats = torch.zeros(4, dtype=torch.int32, device="cuda")
    img, hits = rc.render(cam, 512, 512, light=LIGHT, shadows=True, bilinear=True,
                          return_hits=True, stats=stats)
    rgb, t, prim, st = tw.render(cam.as_array(), 512, 512, LIGHT, shadows=True)
    np.testing.assert_array_equal(hits["prim"].cpu().numpy(), prim)
    np.testing.assert_array_equal(_u32(img.cpu().numpy()), _u32(rgb))
    s = stats.cpu().numpy()
    assert list(s[:3]) == list(st)


@pytest.mark.parametrize("primary,shadow,lane_stack,super_tile,budget", [
    ("lane", "lane", 32, 0, 256), ("lane", "packet", 2, 3, 256), ("packet", "lane", 1, 1, 256),
    ("lane", "lane", 4, 0, 256), ("hybrid", "hybrid", 32, 0, 3), ("hybrid", "lane", 2, 0, 1),
    ("hybrid", "hybrid", 32, 0, 40), ("auto", "lane", 32, 0, 256), ("quad", "lane", 32, 0, 256),
    ("quad", "lane", 32, 2, 256), ("quad", "lane", 32, 5, 256)])
def test_tuning_choices_give_identical_results(c1, twin, primary, shadow, lane_stack, super_tile,
                                               budget):
    """Every traversal / stack / block-mapping / budget choice returns the twin's hits and RGB.
    Stacks of 1-4 entries overflow on most rays, which the waves then trace again with the packet
    traversal; budgets of 1-40 packet steps send most waves on to the per-lane walk."""
    import rendertoy_amd as ren
    meshes, tw, rc = c1
    cam = camera()
    rgb, t, prim, st = tw.render(cam.as_array(), 512, 512, LIGHT, shadows=True)
    ref = tw.trace_camera(cam.as_array(), 512, 512)
    with ren.tuning(primary=primary, shadow=shadow, lane_stack=lane_stack, super_tile=super_tile,
                    packet_budget=budget):
        stats = torch.zeros(4, dtype=torch.int32, device="cuda")
        img, hits = rc.render(cam, 512, 512, light=LIGHT, shadows=True, bilinear=True,
                              return_hits=True, stats=stats)
        gc = {k: v.cpu().numpy() for k, v in rc.cast_camera(cam, 512, 512).items()}
    np.testing.assert_array_equal(hits["prim"].cpu().numpy(), prim)
    np.testing.assert_array_equal(_u32(img.cpu().numpy()), _u32(rgb))
    assert list(stats.cpu().numpy()[:3]) == list(st)
    assert_trace_equal(gc, ref)


@pytest.mark.parametrize("W,H", [(512, 512), (333, 257), (9, 17), (1, 1), (8, 72)])
def test_render_with_and_without_counts(c1, W, H, monkeypatch):
    """rt_render's quad-lane kernel runs one wave per workgroup for a frame without counts and for a
    frame with counts and a workspace (partial counts per slot, summed after), and as 4-wave
    workgroups that sum the counts in LDS for a frame with counts and no workspace: all three give
    the twin's image, hits and counts at square, odd and tiny frame sizes, and the camera kernel (one
    wave per workgroup) its hits."""
    import rendertoy_amd as ren
    meshes, tw, rc = c1
    cam = camera(aspect=W / H)
    rgb, t, prim, st = tw.render(cam.as_array(), W, H, LIGHT, shadows=True)
    ref = tw.trace_camera(cam.as_array(), W, H)
    with ren.tuning(primary="quad"):
        img0, hits0 = rc.render(cam, W, H, light=LIGHT, shadows=True, bilinear=True, return_hits=True)
        stats1 = torch.zeros(4, dtype=torch.int32, device="cuda")
        img1, hits1 = rc.render(cam, W, H, light=LIGHT, shadows=True, bilinear=True, return_hits=True,
                                stats=stats1)
        gc = {k: v.cpu().numpy() for k, v in rc.cast_camera(cam, W, H).items()}
        monkeypatch.setattr(rc, "_render_workspace", lambda n: torch.empty(0, dtype=torch.uint8, device="cuda"))
        stats2 = torch.zeros(4, dtype=torch.int32, device="cuda")
        img2, hits2 = rc.render(cam, W, H, light=LIGHT, shadows=True, bilinear=True, return_hits=True,
                                stats=stats2)
    for img, hits in ((img0, hits0), (img1, hits1), (img2, hits2)):
        np.testing.assert_array_equal(hits["prim"].cpu().numpy(), prim)
        np.testing.assert_array_equal(_u32(hits["t"].cpu().numpy()), _u32(np.where(prim >= 0, t, np.float32(np.inf))))
        np.testing.assert_array_equal(_u32(img.cpu().numpy()), _u32(rgb))
    for stats in (stats1, stats2):
        assert list(stats.cpu().numpy()) == [*st, 0]
    assert_trace_equal(gc, ref)


def test_multi_mesh_nonindexed_int64_and_empty(ren, twin):
    from rendertoy_amd import scenes
    a = scenes.icosphere(8, 11)
    b = scenes.icosphere(4, 12)
    b["pos"] = (b["pos"] * 0.5 + 0.3).astype(np.float32)
    empty = {"pos": np.zeros((3, 3), np.float32), "idx": np.zeros((0, 3), np.int32),
             "nrm": None, "uv": None}
    meshes = [a, empty, b]
    tw = twin_from_meshes(meshes)
    rc_i = gpu_from_meshes(meshes, indexed=True)
    rc_n = gpu_from_meshes(meshes, indexed=False)
    assert_build_equal(rc_i, tw)
    o, d = random_rays(50_000, 13)
    ti = tw.trace_closest(o, d)
    gi = gpu_trace(rc_i, o, d)
    gn = gpu_trace(rc_n, o, d)
    assert_trace_equal(gi, ti)
    np.testing.assert_array_equal(gn["prim"], gi["prim"])
    # (mesh, index) from the global primitive id
    offs = np.array([0, a["idx"].shape[0], a["idx"].shape[0], a["idx"].shape[0] + b["idx"].shape[0]])
    hit = gi["prim"] >= 0
    mesh = np.searchsorted(offs, gi["prim"], side="right") - 1
    mesh = np.where(mesh == 1, 2, mesh)  # mesh 1 is empty: its offset equals mesh 2's
    np.testing.assert_array_equal(gi["mesh"][hit], mesh[hit])
    np.testing.assert_array_equal(gi["index"][hit], (gi["prim"] - offs[mesh])[hit])
    # int64 index buffer (load_obj style) builds the same scene
    m64 = ren.mesh_from_arrays(a["pos"], None, a["nrm"], a["uv"])
    m64.indices = ren.create_buffer_from(a["idx"].reshape(-1).astype(np.int64))
    rc64 = ren.Raycaster([m64])
    tw_a = twin_from_meshes([a])
    assert_build_equal(rc64, tw_a)


def test_tiny_scenes(ren, twin):
    # empty scene: every ray misses
    rc0 = ren.Raycaster([])
    o, d = random_rays(1000, 3)
    g = gpu_trace(rc0, o, d)
    assert (g["prim"] == -1).all() and np.isinf(g["t"]).all()
    # one triangle: root is a leaf
    tri = {"pos": np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32),
           "idx": np.array([[0, 1, 2]], np.int32), "nrm": None, "uv": None}
    tw = twin_from_meshes([tri])
    rc1 = gpu_from_meshes([tri])
    assert rc1.root == ~0
    assert_build_equal(rc1, tw)
    o = np.array([[0, 0, 3], [0, 0, -3], [5, 5, 3], [0, -1, 3]], np.float32)
    d = np.array([[0, 0, -1], [0, 0, 1], [0, 0, -1], [0, 0, -1]], np.float32)
    g = gpu_trace(rc1, o, d)
    assert list(g["prim"]) == [0, 0, -1, 0]  # front, back face (two-sided), miss, on the edge
    np.testing.assert_allclose(g["t"][:2], [3.0, 3.0])
    # two triangles sharing an edge hit at the same t: the smaller primitive id wins
    quad = {"pos": np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32),
            "idx": np.array([[0, 1, 2], [0, 2, 3]], np.int32), "nrm": None, "uv": None}
    rcq = gpu_from_meshes([quad])
    g = gpu_trace(rcq, np.array([[0, 0, 2]], np.float32), np.array([[0, 0, -1]], np.float32))
    assert g["prim"][0] == 0


@pytest.fixture(scope="module")
def c2(ren):
    meshes, tex = c2_scene()
    tw = twin_from_meshes(meshes)
    rc = gpu_from_meshes(meshes)
    torch.cuda.synchronize()
    return meshes, tex, tw, rc


def test_build_c2_matches_twin_bytes(c2):
    meshes, tex, tw, rc = c2
    assert_build_equal(rc, tw)


def test_render_c2_1080p_matches_twin_and_fp64(c2, twin):
    meshes, tex, tw, rc = c2
    W, H = 1920, 1080
    cam = camera(aspect=W / H)
    texg = torch.from_numpy(tex).cuda()
    stats = torch.zeros(4, dtype=torch.int32, device="cuda")
    img, hits = rc.render(cam, W, H, light=LIGHT, textures=[texg], mesh_textures=[0, 0],
                          shadows=True, bilinear=True, return_hits=True, stats=stats)
    rgb, t, prim, st = tw.render(cam.as_array(), W, H, LIGHT, shadows=True, bilinear=True,
                                 textures=[tex], mesh_tex=[0, 0])
    np.testing.assert_array_equal(hits["prim"].cpu().numpy(), prim)
    np.testing.assert_array_equal(_u32(hits["t"].cpu().numpy()), _u32(t))
    np.testing.assert_array_equal(_u32(img.cpu().numpy()), _u32(rgb))
    assert list(stats.cpu().numpy()[:3]) == list(st)
    # the camera kernel gives the same hits as the render's primary pass and the twin, with the
    # 64-ray packets and with the quad-lane 16-ray packets
    import rendertoy_amd as ren
    tref = tw.trace_camera(cam.as_array(), W, H)
    for mode in ("packet", "quad"):
        with ren.tuning(primary=mode):
            gc = {k: v.cpu().numpy() for k, v in rc.cast_camera(cam, W, H).items()}
        np.testing.assert_array_equal(gc["prim"], prim)
        assert_trace_equal(gc, tref)
    # primary IDs and t against FP64 on the full frame
    o, d = twin.twin_generate_rays(cam.as_array(), W, H)
    fp64_report("C2 primary 1920x1080", prim, t, tw.closest64(o, d))
    # shaded RGB against the FP64 shading of the FP64 hits: within 1/255 off flagged pixels
    from rendertoy_amd import scenes
    pos, nrm, uv, mo = scenes.triangle_arrays(meshes)
    rgb64, flags = twin.shade64(pos, nrm, uv, mo, [1, 1], o, d, LIGHT, textures=[tex],
                                mesh_tex=[0, 0], scene=tw)
    err = np.abs(img.cpu().numpy().reshape(-1, 3) - rgb64).max(1)
    print(f"C2 render 1920x1080 vs FP64: rgb_err_max_unflagged={err[flags == 0].max():.3e} "
          f"rgb_err_mean={err[flags == 0].mean():.3e} flagged_pixels={int((flags != 0).sum())} "
          f"over_1/255_flagged={int(((err > 1 / 255) & (flags != 0)).sum())}")
    assert (err[flags == 0] <= 1 / 255).all()


@pytest.mark.parametrize("scene", ["c1", "c2"])
def test_ploc_build_matches_twin_bytes_and_lbvh_hits(ren, scene):
    """The PLOC builder (grid iterations while more than 4096 clusters remain, then one
    workgroup) against the twin's PLOC, byte for byte; its tree gives the LBVH tree's hits."""
    meshes = c1_scene() if scene == "c1" else c2_scene()[0]
    tw = twin_from_meshes(meshes, builder="ploc")
    with ren.tuning(builder="ploc"):
        rc = gpu_from_meshes(meshes)
    torch.cuda.synchronize()
    assert_build_equal(rc, tw)
    lb = gpu_from_meshes(meshes)
    W, H = (512, 512) if scene == "c1" else (960, 540)
    cam = camera(aspect=W / H)
    for mode in ("quad", "packet", "lane"):
        with ren.tuning(primary=mode):
            a = {k: v.cpu().numpy() for k, v in rc.cast_camera(cam, W, H).items()}
            b = {k: v.cpu().numpy() for k, v in lb.cast_camera(cam, W, H).items()}
        np.testing.assert_array_equal(a["prim"], b["prim"])
        np.testing.assert_array_equal(_u32(a["t"]), _u32(b["t"]))
    o, d = random_rays(20000, seed=11)
    np.testing.assert_array_equal(gpu_trace(rc, o, d)["prim"], gpu_trace(lb, o, d)["prim"])


def _edge_geometry_scene():
    """Random triangles plus degenerate ones (points, collinear, a repeated vertex), exact duplicates
    (equal t: the smaller primitive wins), triangles scaled to 1e30, and triangles with a NaN or an
    infinite coordinate."""
    rng = np.random.default_rng(3)
    base = rng.uniform(-1, 1, (200, 3, 3)).astype(np.float32)
    p = rng.uniform(-1, 1, (20, 3)).astype(np.float32)
    nanp = base[:4].copy()
    nanp[:, 1, 0] = np.nan
    infp = base[4:8].copy()
    infp[:, 2, 1] = np.inf
    tris = [base, np.stack([p, p, p], 1), np.stack([p, p + 0.1, p + 0.2], 1).astype(np.float32),
            np.stack([p, p, p + np.float32([0.3, 0, 0])], 1), base[:30].copy(), base[:5] * np.float32(1e30),
            nanp, infp]
    pos = np.concatenate(tris).reshape(-1, 3)
    return [{"pos": pos, "idx": np.arange(len(pos), dtype=np.int32).reshape(-1, 3), "nrm": None, "uv": None}]


@pytest.mark.parametrize("builder", ["lbvh", "ploc"])
def test_edge_geometry(ren, builder):
    """Degenerate, duplicated, huge and non-finite triangles: device == twin bytes, and every ray
    (random, and axis-aligned with zero direction components) hits what FP32 brute force hits."""
    meshes = _edge_geometry_scene()
    tw = twin_from_meshes(meshes, builder=builder)
    with ren.tuning(builder=builder):
        rc = gpu_from_meshes(meshes)
    torch.cuda.synchronize()
    assert_build_equal(rc, tw)
    o, d = random_rays(5000, seed=1)
    rng = np.random.default_rng(4)
    ao = rng.uniform(-1, 1, (600, 3)).astype(np.float32)
    ad = np.zeros((600, 3), np.float32)
    ad[np.arange(600), np.arange(600) % 3] = np.where(np.arange(600) % 2 == 0, 1.0, -1.0)
    ao[np.arange(600), np.arange(600) % 3] = np.where(np.arange(600) % 2 == 0, -3.0, 3.0)
    o, d = np.concatenate([o, ao]), np.concatenate([d, ad])
    g = gpu_trace(rc, o, d)
    t = tw.trace_closest(o, d)
    assert_trace_equal(g, t)
    _, bp = tw.brute_closest32(o, d)
    np.testing.assert_array_equal(g["prim"], bp)
    for mode in ("quad", "packet", "lane"):
        with ren.tuning(primary=mode):
            cam = camera(aspect=1.0)
            a = rc.cast_camera(cam, 96, 96)["prim"].cpu().numpy()
        np.testing.assert_array_equal(a, tw.trace_camera(cam.as_array(), 96, 96)["prim"])


@pytest.mark.parametrize("primary", ["quad", "packet", "lane", "auto"])
def test_render_options(ren, primary):
    """Render flags and parameters against the twin, bit for bit: shadows off and on, point and
    bilinear sampling, ambient/background/shadow_tmin values, odd image sizes (97x61, not a
    multiple of any tile), 1x1 and 3x5 textures and a mesh without a texture (-1), a camera
    whose frame is not coherent (close up and oblique)."""
    from rendertoy_amd import scenes
    a, b = scenes.icosphere(16, 31), scenes.icosphere(8, 32)
    b["pos"] = (b["pos"] * 0.6 + np.float32([0.5, 0.2, 0.1])).astype(np.float32)
    g = scenes.ground_plane()
    meshes = [a, b, g]
    tw = twin_from_meshes(meshes)
    rc = gpu_from_meshes(meshes)
    t1 = np.array([[[200, 10, 30, 255]]], np.uint8)
    t2 = np.random.default_rng(9).integers(0, 256, (5, 3, 4), dtype=np.uint8)
    texs = [t1, t2]
    mt = [1, 0, -1]
    cases = [(97, 61, (0.0, 0.0, 3.0), True, True, 0.2, (0.0, 0.0, 0.0), 1e-4),
             (97, 61, (0.0, 0.0, 3.0), False, False, 0.35, (0.1, 0.2, 0.3), 1e-4),
             (64, 48, (0.9, 0.4, 0.7), True, False, 0.0, (1.0, 0.5, 0.25), 1e-3)]
    with ren.tuning(primary=primary):
        for W, H, eye, shadows, bilinear, ambient, bg, stmin in cases:
            cam = camera(eye=eye, aspect=W / H)
            stats = torch.zeros(4, dtype=torch.int32, device="cuda")
            img, hits = rc.render(cam, W, H, light=LIGHT, textures=[torch.from_numpy(x).cuda() for x in texs],
                                  mesh_textures=mt, shadows=shadows, bilinear=bilinear, ambient=ambient,
                                  background=bg, shadow_tmin=stmin, return_hits=True, stats=stats)
            rgb, t, prim, st = tw.render(cam.as_array(), W, H, LIGHT, ambient=ambient, background=bg,
                                         shadow_tmin=stmin, shadows=shadows, bilinear=bilinear,
                                         textures=texs, mesh_tex=mt)
            np.testing.assert_array_equal(hits["prim"].cpu().numpy(), prim)
            np.testing.assert_array_equal(_u32(img.cpu().numpy()), _u32(rgb))
            assert list(stats.cpu().numpy()[:3]) == list(st)


def _chain_scene(n, base=1.25):
    """n thin triangles at exponentially spaced x: PLOC merges them as a chain (depth ~ n)."""
    x = base ** np.arange(n)
    x = x / x.max() * 4
    z, s = np.zeros(n), np.full(n, 1e-3)
    tri = np.stack([np.stack([x, z, z], 1), np.stack([x, s, z], 1), np.stack([x, z, s], 1)], 1)
    return [{"pos": tri.astype(np.float32).reshape(-1, 3), "idx": np.arange(3 * n, dtype=np.int32).reshape(-1, 3),
             "nrm": None, "uv": None}]


@pytest.mark.parametrize("n", [60, 200])
def test_ploc_depth_guard(ren, n):
    """A PLOC tree deeper than the 62 internal levels the traversal stacks are sized for (200
    chained triangles: 112) is replaced by the LBVH; a shallower chain (60) stays PLOC. Device ==
    twin bytes either way, and rays along the chain hit what brute force hits."""
    meshes = _chain_scene(n)
    tw = twin_from_meshes(meshes, builder="ploc")
    with ren.tuning(builder="ploc"):
        rc = gpu_from_meshes(meshes)
    torch.cuda.synchronize()
    assert_build_equal(rc, tw)
    lb = twin_from_meshes(meshes, builder="lbvh")
    same = np.array_equal(_u32(tw.nodes4), _u32(lb.nodes4))
    assert same == (n == 200)
    k = 256
    o = np.stack([np.linspace(-1, 4.5, k), np.full(k, 2e-4), np.full(k, 2e-4)], 1).astype(np.float32)
    d = np.tile(np.float32([1, 0, 0]), (k, 1))
    o = np.concatenate([o, o * np.float32([1, 1, 1]) + np.float32([0, 6e-4, 6e-4])])
    d = np.concatenate([d, -d])
    _, bp = tw.brute_closest32(o, d)
    np.testing.assert_array_equal(gpu_trace(rc, o, d)["prim"], bp)


def test_ploc_tiny_and_multi_mesh_scenes(ren):
    """PLOC on 2..9 triangles (root a leaf up to 4), duplicated triangles (equal boxes and Morton
    keys) and a multi-mesh scene with an empty mesh: device == twin bytes, hits == LBVH's."""
    from rendertoy_amd import scenes
    rng = np.random.default_rng(21)
    cases = []
    for n in (2, 3, 4, 5, 9):
        pos = rng.uniform(-1, 1, (3 * n, 3)).astype(np.float32)
        cases.append([{"pos": pos, "idx": np.arange(3 * n, dtype=np.int32).reshape(-1, 3),
                       "nrm": None, "uv": None}])
    base = rng.uniform(-1, 1, (30, 3, 3)).astype(np.float32)
    dup = np.concatenate([base, base, base[:5]]).reshape(-1, 3)
    cases.append([{"pos": dup, "idx": np.arange(len(dup), dtype=np.int32).reshape(-1, 3),
                   "nrm": None, "uv": None}])
    a, b = scenes.icosphere(8, 11), scenes.icosphere(4, 12)
    b["pos"] = (b["pos"] * 0.5 + 0.3).astype(np.float32)
    empty = {"pos": np.zeros((3, 3), np.float32), "idx": np.zeros((0, 3), np.int32), "nrm": None, "uv": None}
    cases.append([a, empty, b])
    o, d = random_rays(20_000, 17)
    for meshes in cases:
        tw = twin_from_meshes(meshes, builder="ploc")
        with ren.tuning(builder="ploc"):
            rc = gpu_from_meshes(meshes)
        assert_build_equal(rc, tw)
        assert_trace_equal(gpu_trace(rc, o, d), tw.trace_closest(o, d))
        np.testing.assert_array_equal(gpu_trace(rc, o, d)["prim"], gpu_trace(gpu_from_meshes(meshes), o, d)["prim"])
