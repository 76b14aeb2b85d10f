#!/usr/bin/env python3
"""Closest-hit ray-casting throughput of one MI355X on the 1.31M-triangle, 1920x1080 workload.

Workload (named in the JSON `config`): seeded displaced icosphere of frequency 256 (1,310,720
triangles, per-vertex normals and UVs) plus a 2-triangle ground plane at y=-0.501 (extent +-4),
camera at (0,0,3) looking at the origin, 45 deg vertical FOV, 1920x1080, one primary ray per pixel.

A step = one launch of the fused ray-generation + closest-hit kernel over the whole frame
(2,073,600 rays; hit records t,u,v,prim written per pixel) with the scene resident in HBM.
value = rays / step time in Mrays/s. Also reported: LBVH build time, the full render
(primary + shadow ray + Lambert + bilinear RGBA8 texture) latency p50/p99 over the same count of
frames, the batched / serving / incoherent-path workloads, and ray_cast on explicit ray buffers
(the frame's rays in camera order, and random rays in and around the sphere), and the mesh
entry points on the sphere mesh (Mesh.compute_normals, MeshVertex pack/unpack).

roofline: algorithmic bytes per launch = per ray 16 B of hit output + 128 B per 4-wide node
record + 48 B per triangle record its ray tests, + 48 B per hit (the winner's record re-read for
the FP64 refinement); the per-ray visit counts come from the CPU restatement's lane-exact
emulation of the kernel's traversal on the identical BVH (oracle/rt_twin.c), divided by the
kernel's mean duration from HIP events on the launch stream. The packet traversals fetch a record
once per packet, so the bytes they actually fetch are also reported (fetched_bytes), and
`traffic` is the HBM traffic per launch measured by rocprofv3 (FETCH_SIZE/WRITE_SIZE, scaled by
the same counters on a copy kernel of known size), read from profiles/k_trace_camera_c2.txt when
that summary was taken of the same kernel on the same workload. The quad-lane kernel reads a
record once per 16-ray packet.
`bound` is "latency": the kernel is set by the dependent node fetches of its slowest packets,
not by HBM; `hbm_frac` = measured HBM traffic / kernel time / 8 TB/s.
cpu_baseline: the same closest-hit traversal in the CPU restatement (C + OpenMP, all host
threads) over full frames for about `--cpu-seconds`.
c1: the CPU-runnable workload (81,920 triangles, 512^2, closest hit + Lambert) on the GPU and on
the host (all threads and 1 thread, median of 5).
accuracy: the rendered frame against the FP64 oracle (IDs, flagged rays by kind, t error, RGB
error on unflagged pixels).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s closest-hit (1 GPU) + p50/p99 render latency, batched vs plain"


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--freq", type=int, default=256, help="icosphere frequency (256 = 1.31M tris)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-vs-plain measurement (256 scenes)")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-reps", type=int, default=5)
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the serving measurement (64 resident scenes, Poisson arrivals)")
    ap.add_argument("--serving-requests", type=int, default=1000, help="requests per arrival rate")
    ap.add_argument("--serving-rates", default="50,100,200,500,1000,2000",
                    help="comma-separated arrival rates (requests/s)")
    ap.add_argument("--no-incoherent", action="store_true",
                    help="skip the incoherent-ray measurement (diffuse paths, 4 spp, 2 bounces)")
    ap.add_argument("--no-ray-cast", action="store_true",
                    help="skip the ray_cast measurement (ray buffers in camera order and random)")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the C1 leg (81,920 triangles, 512^2, GPU and CPU at all / 1 thread)")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the FP64 accuracy report of the rendered frame")
    ap.add_argument("--no-mesh", action="store_true",
                    help="skip the mesh entry points (vertex normals, record pack/unpack) on the sphere")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="first run 3 elementwise int32 adds over this many MiB (a kernel of "
                         "known read/write traffic that scales the FETCH_SIZE/WRITE_SIZE "
                         "counters of a profile)")
    return ap.parse_args()


PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "k_trace_camera_c2.txt")


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary, if it was
    taken on this workload and kernel; None otherwise."""
    if not os.path.exists(PROFILE):
        return None
    vals = {}
    with open(PROFILE) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    try:
        if vals.get("workload") != workload or vals.get("kernel") != kernel:
            return None
        return int(float(vals["hbm_bytes_per_launch"]))
    except (KeyError, ValueError):
        return None


def batched_workload(n_scenes, ren, scenes, torch):
    """The batched workload (scenes.batch_scene: scene i from the seed 3000 + i, 1k-500k
    triangles, random transform, image 128^2..1024^2); camera (0,0,3) -> origin, 45 deg; Lambert
    + one shadow ray per lit hit, albedo 1."""
    from rendertoy_amd.batch import RenderRequest
    reqs, tris = [], 0
    for i in range(n_scenes):
        d, res = scenes.batch_scene(i)
        rc = ren.Raycaster([ren.mesh_from_arrays(d["pos"], d["idx"], d["nrm"], d["uv"])])
        cam = ren.Camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.pi / 4, 1.0)
        reqs.append(RenderRequest(rc, cam, res, res))
        tris += rc.n_tris
    torch.cuda.synchronize()
    return reqs, tris


def measure_batched(args, ren, scenes, torch):
    """256 plain rt_render launches back to back vs the same jobs in one graph-captured
    rt_render_batch launch (autotuned launch shape). Latency of a set of requests submitted
    together = time until the last image is done."""
    from rendertoy_amd.batch import BatchRender, autotune, time_launches
    g0 = time.time()
    reqs, tris = batched_workload(args.batch_scenes, ren, scenes, torch)
    gen_s = time.time() - g0
    dev = torch.device("cuda")
    # plain: one rt_render per job, each into its own image, all on one stream
    plain_imgs = [torch.empty((r.height, r.width, 3), dtype=torch.float32, device=dev) for r in reqs]
    plain_stats = torch.zeros((len(reqs), 4), dtype=torch.int32, device=dev)

    def plain():
        for k, r in enumerate(reqs):
            r.raycaster.render(r.camera, r.width, r.height, out=plain_imgs[k], stats=plain_stats[k])
    plain_ms = time_launches(plain, args.batch_reps)
    # the same 256 plain launches replayed from one captured graph: no host launch overhead
    plain()
    torch.cuda.synchronize()
    pg = torch.cuda.CUDAGraph()
    with torch.cuda.graph(pg):
        plain()
    plain_graph_ms = time_launches(pg.replay, args.batch_reps)
    b = BatchRender(reqs)
    best, table = autotune(b, reps=2)
    b.capture()
    batch_ms = time_launches(b.replay, args.batch_reps)
    torch.cuda.synchronize()
    same = all(torch.equal(x.view(torch.int32), y.view(torch.int32)) for x, y in zip(b.images, plain_imgs))
    same = same and torch.equal(b.stats, plain_stats)
    st = b.stats.cpu().numpy()
    rays = int(b.rays_primary + st[:, 1].sum())
    p_med, b_med = float(np.median(plain_ms)), float(np.median(batch_ms))
    return {
        "what": f"{len(reqs)} scenes ({tris} triangles), 128^2..1024^2 images, primary + shadow rays",
        "scenes": len(reqs), "triangles": int(tris), "primary_rays": int(b.rays_primary),
        "rays": rays,
        "plain_ms_p50": round(p_med, 3), "plain_ms_p99": round(percentile(plain_ms, 99), 3),
        "plain_graph_ms_p50": round(float(np.median(plain_graph_ms)), 3),
        "speedup_vs_plain_graph": round(float(np.median(plain_graph_ms)) / b_med, 3),
        "batched_ms_p50": round(b_med, 3), "batched_ms_p99": round(percentile(batch_ms, 99), 3),
        "plain_mrays_per_s": round(rays / (p_med * 1e-3) / 1e6, 2),
        "batched_mrays_per_s": round(rays / (b_med * 1e-3) / 1e6, 2),
        "speedup": round(p_med / b_med, 3),
        "batched_equals_plain": bool(same),
        "autotuned": {"waves_per_cu": best.waves_per_cu, "tiles_per_grab": best.tiles_per_grab,
                      "primary": best.primary, "shadow": best.shadow},
        "autotune_ms": {c.key(): round(ms, 3) for c, ms in table},
        "setup_s": round(gen_s, 1),
    }


def measure_ray_cast(args, rc, cam, W, H, torch):
    """Raycaster.ray_cast's path (rt_trace_closest / rt_trace_any) on explicit ray buffers of the
    bench scene: the frame's primary rays in camera order, and W*H rays with origins uniform in
    [-0.6, 0.6]^3 (around and inside the displaced sphere) and uniformly random directions."""
    from rendertoy_amd.batch import time_launches
    n = W * H
    cr = rc.generate_rays(cam, W, H)
    cam_rays = {k: cr[k] for k in ("ox", "oy", "oz", "dx", "dy", "dz")}
    g = torch.Generator(device="cuda").manual_seed(7)
    o = (torch.rand((n, 3), device="cuda", generator=g) * 2 - 1) * 0.6
    d = torch.randn((n, 3), device="cuda", generator=g)
    d = d / d.norm(dim=1, keepdim=True)
    rnd_rays = {"ox": o[:, 0].contiguous(), "oy": o[:, 1].contiguous(), "oz": o[:, 2].contiguous(),
                "dx": d[:, 0].contiguous(), "dy": d[:, 1].contiguous(), "dz": d[:, 2].contiguous()}
    out = rc.trace_soa(cam_rays)
    res = {"what": f"{n} rays per launch: the frame's primary rays as a ray buffer, and random rays "
                   "(origins in [-0.6,0.6]^3, uniform directions)"}
    for name, rays in (("camera_order", cam_rays), ("random", rnd_rays)):
        ms = float(np.median(time_launches(lambda: rc.trace_soa(rays, out=out), args.batch_reps * 2)))
        ms_any = float(np.median(time_launches(lambda: rc.occluded(rays), args.batch_reps * 2)))
        res[name] = {"closest_ms": round(ms, 3), "closest_mrays_per_s": round(n / ms / 1e3, 1),
                     "any_hit_ms": round(ms_any, 3), "any_hit_mrays_per_s": round(n / ms_any / 1e3, 1),
                     "hits": int((out["prim"] >= 0).sum())}
    return res


def measure_mesh(args, ren, sphere, torch):
    """The mesh entry points on the bench sphere (655,362 vertices, 1,310,720 indexed triangles).
    normals: Mesh.compute_normals (rt_mesh_compute_normals) into a dense float32 [V, 3] tensor and
    in place into the MeshVertex records; algorithmic bytes 12 B of indices per triangle + 12 B of
    positions read and 12 B of normals written per vertex; checked bit-equal against the scene
    generator's numpy normals. pack / unpack: MeshVertex records from / to dense P, N, C tensors
    (rt_mesh_pack / rt_mesh_unpack); algorithmic bytes 32 B of dense attributes + 80 B of records
    per vertex (pack writes whole records, unpack reads the P, N, C fields: 48 B of each record)."""
    from rendertoy_amd.batch import time_launches
    dev = torch.device("cuda")
    nv, nt = len(sphere["pos"]), len(sphere["idx"])
    reps = args.batch_reps * 2
    m = ren.mesh_from_arrays(sphere["pos"], sphere["idx"], None, sphere["uv"])
    out = torch.empty((nv, 3), dtype=torch.float32, device=dev)
    ren.compute_normals(m, out=out)
    same = bool(np.array_equal(out.cpu().numpy().view(np.uint32), sphere["nrm"].view(np.uint32)))
    ms = float(np.median(time_launches(lambda: ren.compute_normals(m, out=out), reps)))
    ms_rec = float(np.median(time_launches(lambda: m.compute_normals(), reps)))
    algo = 12 * nt + 24 * nv
    res = {"normals": {
        "what": f"area-weighted vertex normals, {nv} vertices, {nt} indexed triangles (FP64 face "
                "normals, stable radix sort of the (vertex, pair) records, per-vertex sums)",
        "dense_ms": round(ms, 4), "records_ms": round(ms_rec, 4), "algorithmic_bytes": algo,
        "algorithmic_gbs": round(algo / (ms * 1e-3) / 1e9, 1), "equals_generator": same}}
    lib = ren._native.load()
    t = {k: torch.from_numpy(sphere[k]).to(dev) for k in ("pos", "nrm", "uv")}
    rec = torch.empty(nv * 80, dtype=torch.uint8, device=dev)
    s = ren._native.stream_handle()
    pack = lambda: lib.rt_mesh_pack(t["pos"].data_ptr(), t["nrm"].data_ptr(), t["uv"].data_ptr(), nv,
                                    rec.data_ptr(), s)
    unpack = lambda: lib.rt_mesh_unpack(rec.data_ptr(), nv, t["pos"].data_ptr(), t["nrm"].data_ptr(),
                                        t["uv"].data_ptr(), s)
    for name, fn, algo in (("pack", pack, (32 + 80) * nv), ("unpack", unpack, (32 + 48) * nv)):
        ms = float(np.median(time_launches(fn, reps)))
        res[name] = {"ms": round(ms, 4), "algorithmic_bytes": algo,
                     "algorithmic_gbs": round(algo / (ms * 1e-3) / 1e9, 1)}
    return res


def measure_incoherent(args, ren, scenes, torch):
    """Diffuse paths at 1920x1080, 4 samples per pixel, primary + 2 cosine-weighted bounces
    (Philox4x32-10 key (0x5EED, frame)) on the 8-sphere cluster (scenes.cluster_meshes); rays/s
    counts the rays actually traced (a path stops when it escapes). Per depth: frames with 0, 1 and
    2 bounces are timed alike, and the rays and time a depth adds give its rate. Reported for the
    workload's view (0,0,3) and for scenes.cluster_view(), which frames one sphere so nearly every
    camera ray hits and bounces."""
    from rendertoy_amd.batch import time_launches
    meshes = scenes.cluster_meshes()
    rc = ren.Raycaster([ren.mesh_from_arrays(m["pos"], m["idx"], m["nrm"], m["uv"]) for m in meshes])
    W, H, spp, bounces = 1920, 1080, 4, 2
    out = torch.empty((H, W), dtype=torch.float32, device="cuda")
    res = {"what": "8 spheres x 250,880 triangles, 1920x1080 x 4 spp, primary + 2 diffuse bounces",
           "triangles": int(rc.n_tris), "nominal_rays_per_frame": W * H * spp * (bounces + 1)}
    views = {"workload": ((0.0, 0.0, 3.0), math.pi / 4, (0.0, 0.0, 0.0)), "framed": scenes.cluster_view()}
    for name, (eye, fov, target) in views.items():
        cam = ren.Camera(eye, target, (0.0, 1.0, 0.0), fov, W / H)
        per_b = {}
        for b in range(bounces + 1):
            counts = torch.zeros(b + 1, dtype=torch.int32, device="cuda")
            frame = [0]

            def step():
                frame[0] += 1
                rc.trace_paths(cam, W, H, spp=spp, bounces=b, frame=frame[0], out=out, counts=counts)
            ms = time_launches(step, max(5, args.steps // 5), warmup=2)
            per_b[b] = (float(np.median(ms)), percentile(ms, 99), counts.cpu().numpy().astype(np.int64))
        med, p99, cnt = per_b[bounces]
        depth_rate = []
        for b in range(bounces + 1):
            dt = per_b[b][0] - (per_b[b - 1][0] if b else 0.0)
            depth_rate.append(round(int(cnt[b]) / (dt * 1e-3) / 1e6, 1) if dt > 0 else None)
        res[name] = {"eye": [round(x, 4) for x in eye], "fov": round(fov, 4),
                     "ms_per_frame_p50": round(med, 3), "ms_per_frame_p99": round(p99, 3),
                     "rays_per_frame": int(cnt.sum()), "rays_by_depth": [int(x) for x in cnt],
                     "mrays_per_s": round(int(cnt.sum()) / (med * 1e-3) / 1e6, 2),
                     "ms_by_bounces": [round(per_b[b][0], 3) for b in range(bounces + 1)],
                     "mrays_per_s_by_depth": depth_rate}
    # the workload's figures at top level as before
    for k in ("ms_per_frame_p50", "ms_per_frame_p99", "rays_per_frame", "rays_by_depth", "mrays_per_s"):
        res[k] = res["workload"][k]
    return res


def serving_pool(n_scenes, scenes):
    """The serving pool (scenes.serving_scene: scene k from the seeds 4000 + k / 4100 + k, 10k-2M
    triangles, 512^2 RGBA8 texture); host data is generated when the scene is first made resident."""
    from rendertoy_amd.serving import SceneSpec

    def make(k):
        d, tex = scenes.serving_scene(k)
        return SceneSpec([d], tex)
    return {k: (lambda k=k: make(k)) for k in range(n_scenes)}


def measure_serving(args, ren, scenes, torch):
    """Requests (scene uniform over 64 resident scenes, eye 3d with d uniform on the sphere, image
    256^2 / 512^2 / 1024^2) arriving as a Poisson process at each rate, served by the continuous-
    batching loop (RenderServer: batches on 4 streams, at most 4 in flight) and by the per-request
    loop (PlainServer: one rt_render per request as it arrives) on the same traces; latency = GPU
    completion - arrival. `saturation`: every request present at t = 0 (throughput = requests /
    makespan, the capacity of each loop), after one untimed run of each loop; the median of 3."""
    from rendertoy_amd import serving
    from rendertoy_amd.batch import time_launches
    g0 = time.time()
    store = serving.SceneStore(serving_pool(64, scenes), budget_bytes=200 << 30)
    for k in range(64):
        store.release(store.acquire(k))
    torch.cuda.synchronize()
    setup_s = time.time() - g0
    tris = sum(r.raycaster.n_tris for r in store.resident.values())
    res = store.acquire(0)
    alone = {}
    for size in (256, 512, 1024):
        cam = serving.camera_for((0.0, 0.0, 3.0))
        ms = time_launches(lambda: res.raycaster.render(cam, size, size, textures=res.binding), 5)
        alone[str(size)] = round(float(np.median(ms)), 3)
    store.release(res)
    srv = serving.RenderServer(store)
    plain = serving.PlainServer(store)
    rates = {}
    for rate in [float(x) for x in args.serving_rates.split(",") if x]:
        row = {}
        for name, server in (("batched", srv), ("plain", plain)):
            reqs = serving.make_requests(args.serving_requests, 64, rate, 4000 + int(rate))
            if name == "batched":
                srv.batches = []
            server.run(reqs)
            sm = serving.latency_summary(reqs)
            if name == "batched":
                sm["mean_batch"] = round(float(np.mean(srv.batches)), 2)
            row[name] = sm
        rates[str(int(rate))] = row
        note(f"serving at {rate:.0f} rps: p99 batched {row['batched'].get('p99_ms')} ms, "
             f"plain {row['plain'].get('p99_ms')} ms")
    # saturation: one untimed run per loop first (the batched loop meets its largest batches here
    # for the first time), then 3 runs; the run with the median throughput is reported
    sat = {}
    for name, server in (("batched", srv), ("plain", plain)):
        runs = []
        for k in range(4):
            reqs = serving.make_requests(args.serving_requests, 64, 1e9, 4999)
            if name == "batched":
                srv.batches = []
            server.run(reqs, realtime=False)
            sm = serving.latency_summary(reqs, realtime=False)
            if name == "batched":
                sm["mean_batch"] = round(float(np.mean(srv.batches)), 2)
            if k:
                runs.append(sm)
        runs.sort(key=lambda x: x["throughput_rps"])
        sat[name] = dict(runs[len(runs) // 2])
        sat[name]["throughput_rps_runs"] = [x["throughput_rps"] for x in runs]
    sat["throughput_ratio"] = round(sat["batched"]["throughput_rps"] / sat["plain"]["throughput_rps"], 3)
    return {"what": "64 resident scenes, 512^2 RGBA8 textures, Poisson arrivals, images 256^2..1024^2; "
                    "continuous batching on 4 streams vs one rt_render per request",
            "triangles_resident": int(tris), "resident_bytes": int(store.bytes),
            "alone_render_ms": alone, "rates_rps": rates, "saturation": sat,
            "store": {"hits": store.hits, "misses": store.misses, "evictions": store.evictions},
            "setup_s": round(setup_s, 1)}


def measure_c1(args, ren, scenes, torch):
    """The CPU-runnable workload: icosphere f=64 (81,920 triangles, seed 1000, positions only),
    512x512, 1 ray per pixel from (0,0,3), 45 deg; closest hit + Lambert (L = normalize(1,1,1),
    ambient 0.2), no shadow, no texture. GPU: rt_render; CPU: the twin's per-ray closest hit
    (same BVH, binary nodes) + the same Lambert in numpy, on all host threads and on 1 thread,
    median of 5 runs each."""
    import oracle
    from rendertoy_amd.batch import time_launches
    d = scenes.icosphere(64, 1000)
    rc = ren.Raycaster([ren.mesh_from_arrays(d["pos"], d["idx"])])
    W = H = 512
    cam = ren.Camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.pi / 4, 1.0)
    light = (np.array([1, 1, 1], np.float32) / np.float32(np.sqrt(3.0))).astype(np.float32)
    img = torch.empty((H, W, 3), dtype=torch.float32, device="cuda")
    ms = time_launches(lambda: rc.render(cam, W, H, light=light, shadows=False, out=img), 50, warmup=5)
    gpu_med = float(np.median(ms))
    pos, nrm, uv, mo = scenes.triangle_arrays([{"pos": d["pos"], "idx": d["idx"], "nrm": None, "uv": None}])
    tw = oracle.TwinScene(pos, nrm, uv, mo, [0], [0])
    o, dd = oracle.twin_generate_rays(cam.as_array(), W, H)
    geo = np.cross(pos[:, 3:6] - pos[:, 0:3], pos[:, 6:9] - pos[:, 0:3]).astype(np.float32)
    geo /= np.linalg.norm(geo, axis=1, keepdims=True)

    def cpu_frame():
        h = tw.trace_closest(o, dd)
        hit = h["prim"] >= 0
        ndl = np.maximum(geo[h["prim"][hit]] @ light, 0)
        rgb = np.zeros((W * H, 3), np.float32)
        rgb[hit] = (0.2 + ndl)[:, None]
        return rgb
    threads = host_threads()
    out = {}
    for nt in (threads, 1):
        oracle.set_threads(nt)
        runs = []
        for _ in range(5):
            c0 = time.perf_counter()
            cpu_frame()
            runs.append(time.perf_counter() - c0)
        out[nt] = float(np.median(runs))
    oracle.set_threads(threads)
    n = W * H
    return {"what": "icosphere f=64 (81,920 triangles), 512x512, closest hit + Lambert, no shadow",
            "gpu_ms_p50": round(gpu_med, 4), "gpu_mrays_per_s": round(n / gpu_med / 1e3, 1),
            "cpu_threads": threads, "cpu_ms_median_of_5": round(out[threads] * 1e3, 2),
            "cpu_mrays_per_s": round(n / out[threads] / 1e6, 2),
            "cpu_1thread_ms_median_of_5": round(out[1] * 1e3, 2),
            "cpu_1thread_mrays_per_s": round(n / out[1] / 1e6, 3),
            "gpu_over_cpu": round(out[threads] * 1e3 / gpu_med, 1)}


def host_threads():
    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    return threads


def accuracy_report(tw, meshes, scenes, oracle, cam, W, H, prim, t, img, light, tex):
    """The FP64 accuracy bar on the rendered C2 frame: primary IDs against the FP64 closest hit
    (counted mismatches, unflagged mismatches, flags by kind), t relative error where the IDs
    agree (max, mean), and the RGB error on unflagged pixels against FP64 shading."""
    pos, nrm, uv, mo = scenes.triangle_arrays(meshes)
    o, d = oracle.twin_generate_rays(cam.as_array(), W, H)
    r64 = tw.closest64(o, d)
    flag = r64["edge"]
    mism = prim != r64["prim"]
    agree = (prim >= 0) & ~mism
    rel = np.abs(t[agree].astype(np.float64) - r64["t"][agree]) / r64["t"][agree]
    rgb64, sflags = oracle.shade64(pos, nrm, uv, mo, [True, True], o, d, light, textures=[tex],
                                   mesh_tex=[0, 0], scene=tw, primary=r64)
    ok = (sflags == 0) & ~mism
    err = np.abs(img.reshape(-1, 3).astype(np.float64) - rgb64)[ok]
    return {"what": f"{W}x{H} render of the workload vs the FP64 oracle",
            "rays": int(len(prim)), "hits": int((prim >= 0).sum()),
            "id_mismatch": int(mism.sum()), "unflagged_mismatch": int((mism & (flag == 0)).sum()),
            "flag_edge": int(((flag & oracle.EDGE) != 0).sum()),
            "flag_depth_tie": int(((flag & oracle.DEPTH_TIE) != 0).sum()),
            "flag_near_miss": int(((flag & oracle.NEAR_MISS) != 0).sum()),
            "t_rel_max": float(rel.max()) if rel.size else 0.0,
            "t_rel_mean": float(rel.mean()) if rel.size else 0.0,
            "rgb_max_err_unflagged": float(err.max()) if err.size else 0.0,
            "rgb_pixels_checked": int(ok.sum()), "shadow_flagged": int(((sflags & 8) != 0).sum())}


_T0 = time.time()


def time_launches_stream(fn, steps, warmup, stream):
    """Per-call milliseconds of fn() by HIP events recorded on `stream`, after `warmup` calls."""
    import torch
    for _ in range(warmup):
        fn()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    for a, b in ev:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    return [a.elapsed_time(b) for a, b in ev]


def note(msg):
    """A progress line on stderr (stdout carries only the result line)."""
    print(f"[bench {time.time() - _T0:6.1f} s] {msg}", file=sys.stderr, flush=True)


def percentile(a, q):
    return float(np.percentile(np.asarray(a, dtype=np.float64), q))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested): the ray-casting path "
              "is measured on a single MI355X; multi-GPU sharding is a design note, not a run.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import rendertoy_amd as ren
    from rendertoy_amd import scenes

    W, H = args.width, args.height
    n_rays = W * H
    if args.copy_calib_mib:
        # one elementwise kernel reading and writing exactly `nb` bytes each (int32 add)
        nb = args.copy_calib_mib << 20
        src = torch.ones(nb // 4, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.add(src, 1, out=dst)
        torch.cuda.synchronize()
        del src, dst
    t0 = time.time()
    sphere = scenes.icosphere(args.freq, 2000)
    plane = scenes.ground_plane()
    tex = scenes.value_noise_rgba8(1024, 2001)
    meshes = [sphere, plane]
    models = [ren.mesh_from_arrays(m["pos"], m["idx"], m["nrm"], m["uv"]) for m in meshes]
    gen_s = time.time() - t0

    # ---- build (timed separately) ----
    builds = []
    rc = None
    for _ in range(3):
        torch.cuda.synchronize()
        b0 = time.perf_counter()
        rc = ren.Raycaster(models)
        torch.cuda.synchronize()
        builds.append((time.perf_counter() - b0) * 1e3)
    n_tris = rc.n_tris
    cam = ren.Camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.pi / 4, W / H)
    dev = torch.device("cuda")
    out = {"t": torch.empty(n_rays, dtype=torch.float32, device=dev),
           "u": torch.empty(n_rays, dtype=torch.float32, device=dev),
           "v": torch.empty(n_rays, dtype=torch.float32, device=dev),
           "prim": torch.empty(n_rays, dtype=torch.int32, device=dev)}

    note(f"scene built: {rc.n_tris} triangles, build {np.median(builds):.2f} ms")
    # ---- closest-hit steps ----
    for _ in range(args.warmup):
        rc.cast_camera(cam, W, H, out=out)
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    ev[0].record(stream)
    for k in range(args.steps):
        rc.cast_camera(cam, W, H, out=out)
        ev[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    step_ms = [ev[k].elapsed_time(ev[k + 1]) for k in range(args.steps)]
    total_ms = ev[0].elapsed_time(ev[-1])
    ms_per_step = total_ms / args.steps
    mrays = n_rays / (ms_per_step * 1e-3) / 1e6
    prim = out["prim"].cpu().numpy()
    hits = int((prim >= 0).sum())

    note(f"closest hit: {mrays:.1f} Mrays/s, {ms_per_step:.4f} ms per step")
    # ---- full render latency (primary + shadow + shade), texture bound once ----
    texg = torch.from_numpy(tex).to(dev)
    binding = ren._raycaster.TextureBinding([texg], [0, 0])
    light = (np.array([1, 1, 1], np.float32) / np.float32(np.sqrt(3.0))).astype(np.float32)
    img = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
    # timed as the reference's callers render: without the optional counts (a render that returns
    # them sums them per 4-wave workgroup, reported as latency_ms_p50_with_counts); the counts of the
    # same frame come from one more call
    stats = torch.zeros(4, dtype=torch.int32, device=dev)
    for _ in range(args.warmup):
        rc.render(cam, W, H, light=light, textures=binding, out=img)
    rev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    for a, b in rev:
        a.record(stream)
        rc.render(cam, W, H, light=light, textures=binding, out=img)
        b.record(stream)
    torch.cuda.synchronize()
    render_ms = [a.elapsed_time(b) for a, b in rev]
    img_counted = torch.empty_like(img)
    counted_ms = time_launches_stream(lambda: rc.render(cam, W, H, light=light, textures=binding, out=img_counted,
                                                        stats=stats), args.steps, args.warmup, stream)
    counted_same = bool(torch.equal(img_counted.view(torch.int32), img.view(torch.int32)))
    _, rhits = rc.render(cam, W, H, light=light, textures=binding, out=img, return_hits=True)
    r_prim, r_t = rhits["prim"].cpu().numpy(), rhits["t"].cpu().numpy()
    r_img = img.cpu().numpy()
    # the same frame with per-lane primary walks (tuning primary="lane"): on this dense scene they
    # beat the default packet primary; the image must be the same bits
    lane_tu = {"primary": "lane", "shadow": "lane"}
    img_lane = torch.empty_like(img)
    lane_ms = time_launches_stream(lambda: rc.render(cam, W, H, light=light, textures=binding, out=img_lane,
                                                     tuning=lane_tu), args.steps, args.warmup, stream)
    lane_same = bool(torch.equal(img_lane.view(torch.int32), img.view(torch.int32)))
    # export of the frame to the reference's BGRA8 render-target format (12 B read, 4 B written per px)
    from rendertoy_amd import image as rt_image
    from rendertoy_amd.batch import time_launches
    bgra = rt_image.to_bgra8(img)
    bgra_ms = float(np.median(time_launches(lambda: rt_image.to_bgra8(img, out=bgra), 20)))
    st = stats.cpu().numpy()
    render_rays = n_rays + int(st[1])

    result = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded displaced icosphere + ground plane, value-noise texture)",
        "config": {
            "workload": f"closest-hit primary rays, displaced icosphere f={args.freq} + ground plane, "
                        f"{W}x{H}, 1 ray/pixel, BVH built on device",
            "tris": n_tris, "width": W, "height": H, "rays_per_step": n_rays,
        },
        "step_ms_p50": round(percentile(step_ms, 50), 4),
        "step_ms_p99": round(percentile(step_ms, 99), 4),
        "primary_hits": hits,
        "build_ms_median": round(float(np.median(builds)), 3),
        "render": {
            "what": "primary closest hit + 1 shadow ray per lit hit + Lambert + bilinear RGBA8",
            "latency_ms_p50": round(percentile(render_ms, 50), 4),
            "latency_ms_p99": round(percentile(render_ms, 99), 4),
            "rays_per_frame": render_rays,
            "shadow_rays": int(st[1]),
            "mrays_per_s_p50": round(render_rays / (percentile(render_ms, 50) * 1e-3) / 1e6, 2),
            "bgra8_export_ms": round(bgra_ms, 4),
            "bgra8_export_gbs": round(16 * n_rays / (bgra_ms * 1e-3) / 1e9, 1),
            "lane_primary": {"tuning": lane_tu, "latency_ms_p50": round(percentile(lane_ms, 50), 4),
                             "latency_ms_p99": round(percentile(lane_ms, 99), 4), "image_equal": lane_same},
            "with_counts": {"latency_ms_p50": round(percentile(counted_ms, 50), 4),
                            "latency_ms_p99": round(percentile(counted_ms, 99), 4), "image_equal": counted_same},
        },
        "wall_s_timed": round(wall, 4),
    }

    note(f"render: p50 {percentile(render_ms, 50):.4f} ms (with counts {percentile(counted_ms, 50):.4f} ms, "
         f"lane primary {percentile(lane_ms, 50):.4f} ms)")
    if not args.no_batched:
        note("batched")
        result["batched"] = measure_batched(args, ren, scenes, torch)
    if not args.no_serving:
        note("serving")
        result["serving"] = measure_serving(args, ren, scenes, torch)
    if not args.no_incoherent:
        note("incoherent")
        result["incoherent"] = measure_incoherent(args, ren, scenes, torch)
    if not args.no_ray_cast:
        note("ray_cast")
        result["ray_cast"] = measure_ray_cast(args, rc, cam, W, H, torch)
    if not args.no_mesh:
        note("mesh")
        result["mesh"] = measure_mesh(args, ren, sphere, torch)
    if not args.no_c1:
        note("c1")
        result["c1"] = measure_c1(args, ren, scenes, torch)

    # ---- roofline and CPU baseline from the CPU restatement (oracle) ----
    if not args.no_roofline or not args.no_cpu_baseline or not args.no_accuracy:
        import oracle
        note("oracle: roofline, accuracy, cpu baseline")
        tw_t0 = time.time()
        pos, nrm, uv, mo = scenes.triangle_arrays(meshes)
        tw = oracle.TwinScene(pos, nrm, uv, mo, [1, 1], [1, 1])
        if not args.no_roofline:
            from rendertoy_amd import _native
            quad = ren.get_tuning()["primary"] == _native.RT_TRAVERSAL_QUAD  # the kernel that ran
            wv = tw.trace_camera(cam.as_array(), W, H, quad=quad)
            twin_hits = int((wv["prim"] >= 0).sum())
            lane_nodes = int(wv["lane_nodes"].sum())
            lane_tris = int(wv["lane_tris"].sum())
            algo_bytes = 16 * n_rays + 128 * lane_nodes + 48 * lane_tris + 48 * twin_hits
            fetched = 16 * n_rays + 128 * int(wv["wave_nodes"].sum()) + 48 * int(wv["wave_tris"].sum()) \
                + 48 * twin_hits
            achieved = algo_bytes / (ms_per_step * 1e-3) / 1e9
            traffic = measured_traffic(result["config"]["workload"],
                                       "k_trace_camera_quad" if quad else "k_trace_camera")
            result["roofline"] = {
                # the kernel is bound by the dependent node fetches of its slowest packets (SQ
                # counters: ~40% issue, ~39% memory wait), not by HBM: hbm_frac is what reaches HBM
                "bound": "latency", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": traffic,
                "hbm_gbs": None if traffic is None else round(traffic / (ms_per_step * 1e-3) / 1e9, 1),
                "hbm_frac": None if traffic is None else round(traffic / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
                "algorithmic_bytes_per_launch": int(algo_bytes),
                "bytes_model": "per ray: 16 B hit + 128 B per 4-wide node record + 48 B per "
                               "triangle record its ray tests; + 48 B per hit (FP64 refinement)",
                "mean_node_records_per_ray": round(lane_nodes / n_rays, 4),
                "mean_triangles_per_ray": round(lane_tris / n_rays, 4),
                "fetched_bytes_per_launch": int(fetched),
                "kernel": "k_trace_camera_quad" if quad else "k_trace_camera",
            }
            result["parity_primary_ids_equal_twin"] = bool(np.array_equal(prim, wv["prim"]))
        if not args.no_accuracy:
            note("accuracy")
            result["accuracy"] = accuracy_report(tw, meshes, scenes, oracle, cam, W, H, r_prim, r_t,
                                                 r_img, light, tex)
        if not args.no_cpu_baseline:
            note("cpu baseline")
            o, d = oracle.twin_generate_rays(cam.as_array(), W, H)
            threads = host_threads()
            oracle.set_threads(threads)
            reps, c0 = 0, time.perf_counter()
            while True:
                tw.trace_closest(o, d)
                reps += 1
                if time.perf_counter() - c0 >= args.cpu_seconds:
                    break
            cpu_s = time.perf_counter() - c0
            result["cpu_baseline"] = {
                "value": round(reps * n_rays / cpu_s / 1e6, 3), "unit": "Mrays/s",
                "cores": threads, "kind": "port",
                "sample": f"{reps} full {W}x{H} frames of primary closest hit (per-ray stack "
                          f"traversal of the same BVH), {cpu_s:.1f} s on {threads} host threads",
            }
        result["oracle_s"] = round(time.time() - tw_t0, 2)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
