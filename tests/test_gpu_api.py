"""GPU tests of the drop-in boundary: the BVH export in the reference's record types, per-call
tuning (two tunings rendering concurrently on two streams), the ray-buffer traversals with their
LDS stacks (including forced stack overflow), float4 textures from create_texture2D, the k[n]
launch count, index buffers of every integer type, and a scene on a side stream."""
import math

import numpy as np
import pytest
import torch

from helpers import twin_from_meshes, gpu_from_meshes, camera, random_rays, soa_tensors

pytestmark = pytest.mark.gpu

LIGHT = (np.array([1.0, 1.0, 1.0], np.float32) / np.float32(np.sqrt(3.0))).astype(np.float32)


def _u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


def scene_meshes(f=10, seed=81):
    from rendertoy_amd import scenes
    a = scenes.icosphere(f, seed)
    b = scenes.icosphere(3, seed + 1)
    b["pos"] = (b["pos"] * 0.3 + np.float32(0.4)).astype(np.float32)
    return [a, scenes.ground_plane(), b]


# ---------------------------------------------------------------------------------------------
# export_ads
# ---------------------------------------------------------------------------------------------

def export_reference(tw, meshes):
    """The export rule restated from the twin's hierarchy: breadth-first numbering, node 0 the root
    with the scene bounds, children of the internal nodes of a level appended as pairs in the order
    of their parents; leaves (count, first slot); triangles in slot order with (index, mesh)."""
    n = tw.n
    mesh_of = np.concatenate([np.full(len(m["idx"]), k, np.int32) for k, m in enumerate(meshes)])
    off = np.concatenate([[0], np.cumsum([len(m["idx"]) for m in meshes])]).astype(np.int64)
    tri = np.zeros((n, 16), np.float32)
    prim = tw.order.astype(np.int64)
    p9 = tw.pos.reshape(-1, 9)[prim]
    for k in range(3):
        tri[:, 4 * k:4 * k + 3] = p9[:, 3 * k:3 * k + 3]
    tri[:, 12] = (prim - off[mesh_of[prim]]).astype(np.int32).view(np.float32)
    tri[:, 13] = mesh_of[prim].view(np.float32)
    recs = [[tw.bounds[0:3], tw.bounds[4:7], None, int(tw.root)]]  # lo, hi, count, ref/elements
    lo, hi = 0, 1
    while lo < hi:
        nxt = hi
        for j in range(lo, hi):
            ref = recs[j][3]
            if ref >= 0:
                r = tw.nodes[ref]
                refs = r[12:14].view(np.int32)
                recs.append([r[[0, 2, 8]], r[[1, 3, 9]], None, int(refs[0])])
                recs.append([r[[4, 6, 10]], r[[5, 7, 11]], None, int(refs[1])])
                recs[j][2], recs[j][3] = 0, nxt
                nxt += 2
            else:
                u = (~np.int64(ref)) & 0xFFFFFFFF
                recs[j][2], recs[j][3] = int(u >> 29) + 1, int(u & 0x1FFFFFFF)
        lo, hi = hi, nxt
    nodes = np.zeros((len(recs), 12), np.float32)
    for j, (a, b, c, e) in enumerate(recs):
        nodes[j, 0:3], nodes[j, 4:7] = a, b
        nodes[j, 8:10] = np.array([c, e], np.int32).view(np.float32)
    return nodes, tri


def check_export_semantics(nodes, tris, meshes, n):
    """Leaves partition the triangles; every triangle lies in each box on its path; (index, mesh)
    name each input triangle once."""
    cnt = nodes[:, 8].view(np.int32)
    el = nodes[:, 9].view(np.int32)
    seen = np.zeros(n, np.int32)
    stack = [(0, np.full(3, -np.inf), np.full(3, np.inf))]
    while stack:
        j, plo, phi = stack.pop()
        blo, bhi = nodes[j, 0:3], nodes[j, 4:7]
        if cnt[j] == 0:
            assert el[j] > j
            stack += [(el[j], blo, bhi), (el[j] + 1, blo, bhi)]
            continue
        assert 1 <= cnt[j] <= 4
        for s in range(el[j], el[j] + cnt[j]):
            v = tris[s, [0, 1, 2, 4, 5, 6, 8, 9, 10]].reshape(3, 3)
            assert (v >= blo).all() and (v <= bhi).all()
            seen[s] += 1
    assert (seen == 1).all()
    ids = tris[:, 13].view(np.int32).astype(np.int64) * (1 << 32) + tris[:, 12].view(np.int32)
    assert len(np.unique(ids)) == n


def test_export_ads_equals_the_twin_export(ren):
    meshes = scene_meshes()
    tw = twin_from_meshes(meshes)
    rc = gpu_from_meshes(meshes)
    nodes, tris = rc.export_ads()
    assert nodes.dtype == ren.BVH_AABB and tris.dtype == ren.BVH_Triangle
    g_nodes = nodes.get().view(np.float32).reshape(-1, 12)
    g_tris = tris.get().view(np.float32).reshape(-1, 16)
    r_nodes, r_tris = export_reference(tw, meshes)
    assert g_nodes.shape == r_nodes.shape
    np.testing.assert_array_equal(_u32(g_nodes), _u32(r_nodes))
    np.testing.assert_array_equal(_u32(g_tris), _u32(r_tris))
    check_export_semantics(g_nodes, g_tris, meshes, tw.n)
    # the record fields as the reference's struct names them
    h = nodes.get()
    assert h["count"][0] == 0 and h["elements"][0] == 1 and len(h) == 2 * (h["count"] > 0).sum() - 1


@pytest.mark.parametrize("n_tris", [0, 1, 3, 5])
def test_export_ads_tiny_and_empty_scenes(ren, n_tris):
    rng = np.random.default_rng(n_tris)
    pos = rng.uniform(-1, 1, (max(3 * n_tris, 3), 3)).astype(np.float32)
    idx = np.arange(3 * n_tris, dtype=np.int32).reshape(-1, 3)
    meshes = [{"pos": pos, "idx": idx, "nrm": None, "uv": None}]
    rc = ren.Raycaster([ren.mesh_from_arrays(pos, idx)])
    nodes, tris = rc.export_ads()
    h = nodes.get()
    if n_tris == 0:
        assert len(h) == 1 and h["count"][0] == 0 and h["elements"][0] == -1
        return
    tw = twin_from_meshes(meshes)
    r_nodes, r_tris = export_reference(tw, meshes)
    np.testing.assert_array_equal(_u32(nodes.get().view(np.float32).reshape(-1, 12)), _u32(r_nodes))
    np.testing.assert_array_equal(_u32(tris.get().view(np.float32).reshape(-1, 16)), _u32(r_tris))


def test_export_needs_the_meshes(ren):
    rc = gpu_from_meshes(scene_meshes(4))
    rc.release_inputs()
    with pytest.raises(RuntimeError, match="release_inputs"):
        rc.export_ads()


# ---------------------------------------------------------------------------------------------
# ray buffers: BINARY / LANE walks with LDS stacks, overflow retraced by the packet traversal
# ---------------------------------------------------------------------------------------------

@pytest.mark.parametrize("query", ["binary", "lane"])
@pytest.mark.parametrize("lane_stack", [1, 2, 5, 32])
def test_ray_buffer_traversals_and_stack_overflow(ren, query, lane_stack):
    from rendertoy_amd import scenes
    meshes = [scenes.icosphere(24, 91), scenes.ground_plane()]
    tw = twin_from_meshes(meshes)
    rc = gpu_from_meshes(meshes)
    o, d = random_rays(30000, 92)
    # rays starting inside the sphere, where many boxes are pending at once
    rng = np.random.default_rng(93)
    o2 = rng.uniform(-0.3, 0.3, (20000, 3)).astype(np.float32)
    d2 = rng.standard_normal((20000, 3))
    d2 = (d2 / np.linalg.norm(d2, axis=1, keepdims=True)).astype(np.float32)
    o, d = np.concatenate([o, o2]), np.concatenate([d, d2])
    tu = {"query": query, "lane_stack": lane_stack}
    g = rc.trace_soa(soa_tensors(o, d), tuning=tu)
    ref = tw.trace_closest(o, d)
    np.testing.assert_array_equal(g["prim"].cpu().numpy(), ref["prim"])
    hit = ref["prim"] >= 0
    np.testing.assert_array_equal(_u32(g["t"].cpu().numpy()[hit]), _u32(ref["t"][hit]))
    np.testing.assert_array_equal(_u32(g["u"].cpu().numpy()[hit]), _u32(ref["u"][hit]))
    occ = rc.occluded(soa_tensors(o, d), tuning=tu).cpu().numpy()
    np.testing.assert_array_equal(occ, tw.trace_any(o, d))
    # tmin / tmax per ray
    t = soa_tensors(o, d)
    t["tmin"] = torch.full((len(o),), 0.5, dtype=torch.float32, device="cuda")
    t["tmax"] = torch.full((len(o),), 2.5, dtype=torch.float32, device="cuda")
    g2 = rc.trace_soa(t, tuning=tu)
    r2 = tw.trace_closest(o, d, tmin=0.5, tmax=2.5)
    np.testing.assert_array_equal(g2["prim"].cpu().numpy(), r2["prim"])


@pytest.mark.parametrize("n_tris", [1, 2, 4])
def test_leaf_root_scenes_through_every_lane_walk(ren, n_tris):
    """At most 4 triangles make one leaf at the root: a per-lane walk parks it at its first step
    with no node left to visit. The queries (LANE and BINARY), the render under every primary
    traversal and the batch refill kernel give the twin's hits and the same image."""
    from rendertoy_amd.batch import BatchRender, BatchConfig, RenderRequest
    rng = np.random.default_rng(700 + n_tris)
    mesh = {"pos": rng.uniform(-0.6, 0.6, (3 * n_tris, 3)).astype(np.float32),
            "idx": np.arange(3 * n_tris, dtype=np.int32).reshape(-1, 3), "nrm": None, "uv": None}
    tw = twin_from_meshes([mesh])
    rc = gpu_from_meshes([mesh])
    assert rc.root < 0
    o, d = random_rays(4000, 701 + n_tris)
    ref = tw.trace_closest(o, d)
    hit = ref["prim"] >= 0
    assert hit.any() and not hit.all()
    for q in ("lane", "binary"):
        g = rc.trace_soa(soa_tensors(o, d), tuning={"query": q})
        np.testing.assert_array_equal(g["prim"].cpu().numpy(), ref["prim"])
        np.testing.assert_array_equal(_u32(g["t"].cpu().numpy()[hit]), _u32(ref["t"][hit]))
        occ = rc.occluded(soa_tensors(o, d), tuning={"query": q}).cpu().numpy()
        np.testing.assert_array_equal(occ, tw.trace_any(o, d))
    W, H = 96, 64
    cam = camera((0.2, 0.3, 2.5), aspect=W / H)
    rgb, t, prim, st = tw.render(cam.as_array(), W, H, LIGHT)
    assert (prim >= 0).any()
    for primary in ("packet", "lane", "auto", "hybrid", "quad"):
        img = rc.render(cam, W, H, light=LIGHT, tuning={"primary": primary, "shadow": "lane"})
        np.testing.assert_array_equal(_u32(img.cpu().numpy()), _u32(rgb), err_msg=primary)
    b = BatchRender([RenderRequest(rc, cam, W, H, light=LIGHT)] * 2, config=BatchConfig(16, 1, "lane", "lane"))
    b.launch()
    torch.cuda.synchronize()
    for im in b.images:
        np.testing.assert_array_equal(_u32(im.cpu().numpy()), _u32(rgb))


@pytest.mark.parametrize("query,lane_stack", [("lane", 1), ("binary", 2), ("lane", 32)])
def test_long_ray_buffers_refill_lanes_and_retrace_overflows(ren, query, lane_stack):
    """A buffer of 1,500,001 rays: more 64-ray groups than the GPU holds waves, so every wave refills
    its lanes from several groups; with 1- and 2-entry stacks most rays overflow, so the waves' LDS
    lists fill, refills pause and the listed rays are traced again with the packet traversal
    between phases. Every hit equals the twin's."""
    from rendertoy_amd import scenes
    meshes = [scenes.icosphere(24, 91), scenes.ground_plane()]
    tw = twin_from_meshes(meshes)
    rc = gpu_from_meshes(meshes)
    n = 1_500_001
    o, d = random_rays(n, 97)
    rng = np.random.default_rng(98)
    k = n // 3  # a third start inside the sphere
    o[:k] = rng.uniform(-0.3, 0.3, (k, 3)).astype(np.float32)
    tu = {"query": query, "lane_stack": lane_stack}
    g = rc.trace_soa(soa_tensors(o, d), tuning=tu)
    ref = tw.trace_closest(o, d)
    np.testing.assert_array_equal(g["prim"].cpu().numpy(), ref["prim"])
    hit = ref["prim"] >= 0
    np.testing.assert_array_equal(_u32(g["t"].cpu().numpy()[hit]), _u32(ref["t"][hit]))
    occ = rc.occluded(soa_tensors(o, d), tuning=tu).cpu().numpy()
    np.testing.assert_array_equal(occ, tw.trace_any(o, d))
    assert hit.mean() > 0.3


# ---------------------------------------------------------------------------------------------
# per-call tuning, concurrently on two streams
# ---------------------------------------------------------------------------------------------

def test_two_tunings_render_concurrently_on_two_streams(ren, twin):
    from rendertoy_amd import scenes
    meshes = [scenes.icosphere(40, 95), scenes.ground_plane()]
    tex = scenes.value_noise_rgba8(128, 96)
    tw = twin_from_meshes(meshes)
    W, H = 320, 240
    cam = camera((0.3, 0.6, 2.4), (0, 0, 0), (0, 1, 0), math.pi / 4, W / H)
    ref, _, ref_prim, _ = tw.render(cam.as_array(), W, H, LIGHT, textures=[tex], mesh_tex=[0, -1])
    default = ren.get_tuning()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    rc1 = gpu_from_meshes(meshes)
    torch.cuda.synchronize()
    rc1 = ren.Raycaster(rc1.models, stream=s1, tuning={"primary": "packet", "shadow": "packet"})
    rc2 = ren.Raycaster(rc1.models, stream=s2, tuning={"primary": "quad", "builder": "ploc"})
    texg = torch.from_numpy(tex).cuda()
    torch.cuda.synchronize()
    outs = []
    for rep in range(4):
        a = rc1.render(cam, W, H, light=LIGHT, textures=[texg], mesh_textures=[0, -1], return_hits=True)
        b = rc2.render(cam, W, H, light=LIGHT, textures=[texg], mesh_textures=[0, -1], return_hits=True,
                       tuning={"shadow": "hybrid", "packet_budget": 3})
        outs.append((a, b))
    torch.cuda.synchronize()
    assert ren.get_tuning() == default  # per-call choices never touch the process default
    for (img1, h1), (img2, h2) in outs:
        np.testing.assert_array_equal(_u32(img1.cpu().numpy()), _u32(ref))
        np.testing.assert_array_equal(_u32(img2.cpu().numpy()), _u32(ref))
        np.testing.assert_array_equal(h1["prim"].cpu().numpy(), ref_prim)
        np.testing.assert_array_equal(h2["prim"].cpu().numpy(), ref_prim)


def test_side_stream_scene_queries_and_batch(ren):
    """Raycaster(stream=s): the build waits for the caller's uploads, and ray_cast, render and a
    BatchRender on s give what the default stream gives."""
    from rendertoy_amd.batch import BatchRender, RenderRequest
    meshes = scene_meshes(16, 97)
    base = gpu_from_meshes(meshes)
    o, d = random_rays(5000, 98)
    rays = ren.create_buffer(len(o), ren.Ray)
    rec = np.zeros(len(o), ren.Ray)
    rv = rec.view(np.float32).reshape(-1, 8)
    rv[:, 0:3], rv[:, 4:7] = o, d
    rays.set(rec)
    cam = camera((0.2, 0.3, 2.5), (0, 0, 0), (0, 1, 0), 0.8, 4 / 3)
    want_hits = base.ray_cast(rays).get()
    want_img = base.render(cam, 160, 120, light=LIGHT).cpu().numpy()
    s = torch.cuda.Stream()
    models = [ren.mesh_from_arrays(m["pos"], m["idx"], m.get("nrm"), m.get("uv")) for m in meshes]
    rc = ren.Raycaster(models, stream=s)  # the uploads above are on the current stream
    hits = rc.ray_cast(rays)
    img = rc.render(cam, 160, 120, light=LIGHT)
    b = BatchRender([RenderRequest(rc, cam, 160, 120, light=LIGHT)], stream=s)
    b.launch()
    torch.cuda.current_stream().wait_stream(s)
    np.testing.assert_array_equal(hits.get().view(np.uint32), want_hits.view(np.uint32))
    np.testing.assert_array_equal(_u32(img.cpu().numpy()), _u32(want_img))
    np.testing.assert_array_equal(_u32(b.images[0].cpu().numpy()), _u32(want_img))


# ---------------------------------------------------------------------------------------------
# reference formats at the boundary
# ---------------------------------------------------------------------------------------------

def unorm8_np(x):
    x = np.asarray(x, np.float32) * np.float32(255.0)
    out = np.where(x > 0, np.rint(np.minimum(x, 255)), 0)
    return np.where(np.isnan(x), 0, out).astype(np.uint8)


def test_float4_textures_from_create_texture2D(ren):
    """create_texture2D(w, h) holds float4 texels (the reference's default, rendering/_core.py:
    557-578): rendering with it equals rendering with its UNORM_INT8 conversion, which equals the
    numpy statement of that rule."""
    from rendertoy_amd import scenes
    meshes = [scenes.icosphere(12, 99), scenes.ground_plane()]
    rc = gpu_from_meshes(meshes)
    rng = np.random.default_rng(100)
    w, h = 37, 23
    texels = rng.uniform(-0.1, 1.1, (h, w, 4)).astype(np.float32)
    texels[0, 0] = [0.5 / 255, 127.5 / 255, np.nan, np.inf]
    mem, desc = ren.create_texture2D(w, h)
    assert mem.dtype == ren.float4
    with ren.mapped(mem) as m:
        m.view(np.float32).reshape(h, w, 4)[:] = texels
    rgba8 = unorm8_np(texels)
    cam = camera((0.4, 0.2, 2.5), (0, 0, 0), (0, 1, 0), 0.8, 1.5)
    img_f = rc.render(cam, 150, 100, light=LIGHT, textures=[(mem, desc)], mesh_textures=[0, 0])
    img_8 = rc.render(cam, 150, 100, light=LIGHT, textures=[torch.from_numpy(rgba8).cuda()],
                      mesh_textures=[0, 0])
    np.testing.assert_array_equal(_u32(img_f.cpu().numpy()), _u32(img_8.cpu().numpy()))
    img_t = rc.render(cam, 150, 100, light=LIGHT, textures=[torch.from_numpy(texels).cuda()],
                      mesh_textures=[0, 0])
    np.testing.assert_array_equal(_u32(img_t.cpu().numpy()), _u32(img_8.cpu().numpy()))
    binding = ren._raycaster.TextureBinding([(mem, desc)], [0, 0])
    img_b = rc.render(cam, 150, 100, light=LIGHT, textures=binding)
    np.testing.assert_array_equal(_u32(img_b.cpu().numpy()), _u32(img_8.cpu().numpy()))


def test_camera_from_reference_matrices(ren, twin):
    """Camera.from_matrices(look_at(...), perspective(...)) traces what the twin traces with the
    same camera vectors, and its rays agree with the FP64 closed form."""
    from rendertoy_amd import scenes
    import oracle
    meshes = [scenes.icosphere(20, 101), scenes.ground_plane()]
    tw = twin_from_meshes(meshes)
    rc = gpu_from_meshes(meshes)
    W, H = 199, 131
    eye, tgt = (0.7, 0.9, 2.1), (0.05, -0.1, 0.0)
    view = ren.look_at(ren.make_float3(*eye), ren.make_float3(*tgt), ren.make_float3(0, 1, 0))
    proj = ren.perspective(0.9, W / H)
    cam = ren.Camera.from_matrices(view, proj)
    g = rc.cast_camera(cam, W, H)
    o, d = oracle.twin_generate_rays(cam.as_array(), W, H)
    ref = tw.trace_closest(o, d)
    np.testing.assert_array_equal(g["prim"].cpu().numpy(), ref["prim"])
    d64 = oracle.camera_dirs_fp64(eye, tgt, (0, 1, 0), 0.9, W, H)
    assert np.abs(d.astype(np.float64) - d64).max() < 2e-6


def test_trace_n_launches_exactly_n_rays(ren):
    from rendertoy_amd import scenes
    meshes = [scenes.icosphere(8, 103)]
    rc = gpu_from_meshes(meshes)
    o, d = random_rays(1000, 104)
    rays = soa_tensors(o, d)
    full = rc.trace_soa(rays)
    hits = {k: torch.full((1000,), -7, dtype=v.dtype, device="cuda") for k, v in full.items()}
    rc.trace[321](rays, hits)
    for k in full:
        np.testing.assert_array_equal(hits[k][:321].cpu().numpy(), full[k][:321].cpu().numpy())
        assert (hits[k][321:] == -7).all()
    small = {"prim": torch.empty(300, dtype=torch.int32, device="cuda")}
    with pytest.raises(ValueError, match="elements"):
        rc.trace[321](rays, small)
    with pytest.raises(ValueError, match="rays requested"):
        rc.trace[1001](rays, None)
    # the reference's shape form: n = prod(shape)
    out = rc.trace[(10, 20)](rays, None)
    np.testing.assert_array_equal(out["prim"].cpu().numpy(), full["prim"][:200].cpu().numpy())


@pytest.mark.parametrize("dtype", [np.uint32, np.int64, np.int16, np.uint16])
def test_index_buffers_of_every_integer_type(ren, dtype):
    from rendertoy_amd import scenes
    d = scenes.icosphere(12, 105)
    ref = gpu_from_meshes([d])
    nv = len(d["pos"])
    if np.iinfo(dtype).max < nv:
        d = scenes.icosphere(5, 105)
        ref = gpu_from_meshes([d])
    idx = d["idx"].astype(dtype)
    m = ren.mesh_from_arrays(d["pos"], None, d["nrm"], d["uv"])
    m.indices = ren.create_buffer_from(idx.reshape(-1))
    rc = ren.Raycaster([m])
    a, b = ref.scene_arrays(), rc.scene_arrays()
    for k in ("nodes", "nodes4", "tris", "shade", "order"):
        assert torch.equal(a[k].view(torch.int32), b[k].view(torch.int32)), k


def test_out_of_range_indices_clamp_the_same_for_every_type(ren):
    rng = np.random.default_rng(106)
    pos = rng.uniform(-1, 1, (50, 3)).astype(np.float32)
    idx = rng.integers(0, 50, (40, 3)).astype(np.int64)
    idx[3, 1] = 70            # past the end
    idx[7, 2] = -4            # negative
    idx[9, 0] = 2 ** 32 + 5   # would wrap to 5 when narrowed naively
    m64 = ren.mesh_from_arrays(pos, None)
    m64.indices = ren.create_buffer_from(idx.reshape(-1))
    clamped = np.where((idx < 0) | (idx >= 50), 49, idx).astype(np.int32)
    ref = ren.Raycaster([ren.mesh_from_arrays(pos, clamped)])
    rc = ren.Raycaster([m64])
    assert torch.equal(ref.scene_arrays()["tris"].view(torch.int32), rc.scene_arrays()["tris"].view(torch.int32))


def test_bad_arguments_raise_before_any_launch(ren):
    from rendertoy_amd import scenes
    meshes = [scenes.icosphere(4, 107), scenes.ground_plane()]
    rc = gpu_from_meshes(meshes)
    o, d = random_rays(100, 108)
    rays = soa_tensors(o, d)
    with pytest.raises(ValueError, match="elements"):
        rc.trace_soa(rays, out={"t": torch.empty(50, dtype=torch.float32, device="cuda")})
    with pytest.raises(TypeError):
        rc.trace_soa(rays, out={"prim": torch.empty(100, dtype=torch.float32, device="cuda")})
    with pytest.raises(RuntimeError, match="HIP device"):
        rc.trace_soa(rays, out={"t": torch.empty(100, dtype=torch.float32)})
    short = dict(rays)
    short["dz"] = short["dz"][:99].clone()
    with pytest.raises(ValueError, match="rays\\['dz'\\]"):
        rc.trace_soa(short)
    with pytest.raises(ValueError, match="mesh_textures"):
        rc.render(camera(), 16, 16, textures=[torch.zeros((4, 4, 4), dtype=torch.uint8, device="cuda")],
                  mesh_textures=[0])
    with pytest.raises(ValueError, match="elements"):
        rc.render(camera(), 16, 16, out=torch.empty((16, 16, 2), dtype=torch.float32, device="cuda"))
    with pytest.raises(ValueError, match="elements"):
        rc.cast_camera(camera(), 16, 16, out={"t": torch.empty(100, dtype=torch.float32, device="cuda")})
    with pytest.raises(RuntimeError, match="tuning"):
        rc.cast_camera(camera(), 16, 16, tuning={"lane_stack": 99})
