"""Render serving: resident scenes under a device-memory budget (LRU) and request loops.

The reference has no serving layer: its only request loop is the interactive frame loop of the
tutorials (tutorials/lesson09_texture_mapping.py:112-143), one frame per iteration. Two loops run
the same request traces here:

- `RenderServer`: continuous batching. While batches run on the GPU the loop admits arrivals,
  makes their scenes resident (building the BVH on a miss, evicting least-recently-used scenes
  that no in-flight batch uses) and launches the requests admitted since the previous launch (at
  most `max_batch`, grouped by scene) as one rt_render_batch launch (up to `plain_max` requests as
  rt_render calls, in arrival order on the first stream), on the least busy of `n_streams`
  equal-priority streams: batches on different streams overlap on the device (one batch's tail
  tiles share the CUs with the next batch's bulk), and a new batch starts as soon as a slot (at
  most `max_inflight` batches in flight) frees up.
- `PlainServer`: the per-request baseline: one rt_render launch per request as it arrives, in
  arrival order on one stream (what a caller of the plain entry point does).

A request's latency is the GPU completion time of its work (HIP event) minus its arrival time.
"""
import collections
import dataclasses
import math
import time
import typing

import numpy as np
import torch

from . import _native as N
from .batch import BatchConfig, BatchSlot, RenderRequest


@dataclasses.dataclass
class SceneSpec:
    """Host-side description of a scene: meshes as dicts of arrays (pos, idx, nrm, uv) and an
    optional RGBA8 texture [H, W, 4] applied to every mesh."""
    meshes: list
    texture: typing.Any = None

    def tris(self) -> int:
        return int(sum(len(m["idx"]) for m in self.meshes))


class Resident:
    """A scene held on the device: its Raycaster, texture and their device bytes."""

    def __init__(self, raycaster, texture, nbytes):
        self.raycaster = raycaster
        self.texture = texture
        self.binding = None
        if texture is not None:
            from ._raycaster import TextureBinding
            self.binding = TextureBinding([texture], [0] * raycaster.n_models)
        self.nbytes = nbytes
        self.pins = 0


def scene_bytes_estimate(n_tris: int, texture=None):
    """(resident, transient) device bytes of a scene of n_tris triangles: its arrays and texture,
    and the build workspace plus input vertices it needs while it is built."""
    lib = N.load()
    import ctypes
    s = N.rt_scene_sizes()
    N.check(lib.rt_scene_sizes_for(max(1, n_tris), ctypes.byref(s)), "rt_scene_sizes_for")
    resident = s.nodes_bytes + s.nodes4_bytes + s.tris_bytes + s.shade_bytes + 8 * n_tris + 64
    tex = 0 if texture is None else int(np.asarray(texture).nbytes)
    return int(resident + tex), int(s.workspace_bytes + 3 * 80 * n_tris)


def build_resident(spec: SceneSpec) -> Resident:
    """Uploads a scene's meshes and texture and builds its BVH on the device."""
    from ._modeling import mesh_from_arrays
    from ._raycaster import Raycaster
    models = [mesh_from_arrays(m["pos"], m["idx"], m.get("nrm"), m.get("uv")) for m in spec.meshes]
    rc = Raycaster(models)
    rc.release_inputs()
    tex = None
    if spec.texture is not None:
        tex = torch.from_numpy(np.ascontiguousarray(spec.texture, np.uint8)).cuda()
    return Resident(rc, tex, rc.device_bytes() + (0 if tex is None else tex.numel()))


class SceneStore:
    """Resident scenes keyed by id, LRU-evicted to keep their device bytes within `budget_bytes`.
    `specs` maps id -> SceneSpec (or a callable returning one). Scenes pinned by an in-flight
    batch are never evicted."""

    def __init__(self, specs, budget_bytes: int, build=None, estimate=None):
        self.specs = specs
        self.budget = int(budget_bytes)
        self.resident = collections.OrderedDict()
        self.bytes = 0
        self.hits = self.misses = self.evictions = 0
        self._build = build or build_resident
        self._estimate = estimate or (lambda spec: scene_bytes_estimate(spec.tris(), spec.texture)[0])

    def _spec(self, sid):
        s = self.specs[sid]
        return s() if callable(s) else s

    def _evict_for(self, need):
        for sid in list(self.resident.keys()):
            if self.bytes + need <= self.budget:
                break
            r = self.resident[sid]
            if r.pins:
                continue
            del self.resident[sid]
            self.bytes -= r.nbytes
            self.evictions += 1
        if self.bytes + need > self.budget:
            raise MemoryError(f"scene of {need} bytes does not fit the budget {self.budget} "
                              f"({self.bytes} bytes pinned by queued batches)")

    def acquire(self, sid) -> Resident:
        """The resident scene `sid` (built on a miss), marked most recently used and pinned."""
        r = self.resident.get(sid)
        if r is not None:
            self.resident.move_to_end(sid)
            self.hits += 1
        else:
            spec = self._spec(sid)
            self._evict_for(self._estimate(spec))
            r = self._build(spec)
            self.resident[sid] = r
            self.bytes += r.nbytes
            self.misses += 1
        r.pins += 1
        return r

    def release(self, r: Resident):
        r.pins -= 1


@dataclasses.dataclass
class Request:
    """One render request: scene id, camera eye (looking at the origin), square image size,
    arrival time in seconds from the start of the run."""
    scene: typing.Any
    eye: tuple
    size: int
    arrival: float = 0.0
    keep_image: bool = False
    # filled in by the server
    latency: float = math.nan
    service: float = math.nan   # GPU time of the request's batch
    batch: int = -1
    image: typing.Any = None


def camera_for(eye):
    """Camera at `eye` looking at the origin, 45 deg, aspect 1, up +y (+z when the view is
    near-vertical): the basis of Camera(eye, origin, up) computed on Python floats in double
    precision and stored as float32 (the serving loop makes one camera per request, and numpy's
    per-call overhead on 3-vectors was ~60 us of host time per camera)."""
    from ._raycaster import Camera
    ex, ey, ez = (float(v) for v in eye)
    n = math.sqrt(ex * ex + ey * ey + ez * ez)
    zx, zy, zz = -ex / n, -ey / n, -ez / n
    ux, uy, uz = (0.0, 0.0, 1.0) if abs(zy) > 0.99 else (0.0, 1.0, 0.0)
    xx, xy, xz = uy * zz - uz * zy, uz * zx - ux * zz, ux * zy - uy * zx  # up x z, normalised
    m = math.sqrt(xx * xx + xy * xy + xz * xz)
    xx, xy, xz = xx / m, xy / m, xz / m
    yx, yy, yz = zy * xz - zz * xy, zz * xx - zx * xz, zx * xy - zy * xx  # z x x
    t = math.tan(math.pi / 8)
    cam = Camera.__new__(Camera)
    cam.eye = np.array((ex, ey, ez), np.float32)
    cam.right = np.array((xx * t, xy * t, xz * t), np.float32)
    cam.up = np.array((yx * t, yy * t, yz * t), np.float32)
    cam.forward = np.array((zx, zy, zz), np.float32)
    cam.fov = math.pi / 4
    cam.aspect = 1.0
    return cam


class RenderServer:
    """Serves Requests in arrival order by continuous batching (see the module docstring)."""

    def __init__(self, store: SceneStore, config: BatchConfig = None, max_batch: int = 128,
                 max_inflight: int = 4, n_streams: int = 4, plain_max: int = 4):
        self.store = store
        # the best launch shape on the saturated serving workload (500 requests over the 64-scene
        # pool): 16 waves per CU, 4 tiles per grab, per-lane primary and shadow walks with refill
        self.config = config or BatchConfig(waves_per_cu=16, tiles_per_grab=4, shadow="lane", primary="lane")
        # at most 128 requests per batch: a backlog goes out as several batches that overlap on
        # the device (saturated serving workload: 12.9k requests/s at 128, 12.3k at 256, 6.8k as
        # one batch of 1000, whose job table the host must build before anything starts)
        self.max_batch = max_batch
        # launches of at most plain_max requests run as rt_render calls in arrival order on the
        # first stream, as the plain loop runs them: at 500-1000 requests/s, batch-kernel launches
        # of 2-4 requests raised the p99 by ~1 ms over plain, these keep it equal
        self.plain_max = plain_max
        self.max_inflight = max(1, min(max_inflight, n_streams))
        # equal (default) priority streams: batches on different streams overlap on the device; one
        # reusable job table and workspace per stream (at most one batch per stream in flight)
        self.streams = [torch.cuda.Stream() for _ in range(max(1, n_streams))]
        self.slots = {id(st): BatchSlot(max_batch, st, self.config) for st in self.streams}
        self.batches = []      # requests per launched batch
        self.batch_log = []    # (requests, GPU start s, GPU end s) per batch, from the run's start

    def _pick_stream(self, inflight):
        busy = collections.Counter(id(f["stream"]) for f in inflight)
        return min(self.streams, key=lambda st: busy[id(st)])

    def _launch(self, reqs, t0_event, stream):
        """Launches the longest prefix of `reqs` whose scenes fit the budget next to the pinned
        ones on `stream`; returns (record or None, requests left over)."""
        residents = []
        for k, r in enumerate(reqs):
            try:
                residents.append(self.store.acquire(r.scene))
            except MemoryError:
                if not residents:
                    return None, reqs
                reqs, rest = reqs[:k], reqs[k:]
                break
        else:
            rest = []
        small = len(reqs) <= self.plain_max
        if not small:
            # the batch's jobs grouped by scene (larger images first): tiles of one scene follow
            # each other in the batch's tile order, so its hierarchy stays in the caches
            order = sorted(range(len(reqs)), key=lambda i: (reqs[i].scene, -reqs[i].size))
            reqs = [reqs[i] for i in order]
            residents = [residents[i] for i in order]
        rr = [RenderRequest(res.raycaster, camera_for(r.eye), r.size, r.size, textures=res.binding)
              for r, res in zip(reqs, residents)]
        # scenes built on the current stream are ready before the batch starts
        stream.wait_stream(torch.cuda.current_stream())
        start = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        events = None
        if small:
            # a few requests: one rt_render each, in arrival order (the batch kernel pays off only
            # for larger batches); each request completes at its own event
            images, events = [], []
            with torch.cuda.stream(stream):
                for r in rr:
                    images.append(r.raycaster.render(r.camera, r.width, r.height, textures=r.textures))
                    events.append(torch.cuda.Event(enable_timing=True))
                    events[-1].record(stream)
            ev = events[-1]
        else:
            images = self.slots[id(stream)].launch(rr, self.config)
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(stream)
        return {"reqs": reqs, "images": images, "event": ev, "events": events, "start": start,
                "residents": residents, "t0": t0_event, "stream": stream}, rest

    def _finish(self, f):
        done_s = f["t0"].elapsed_time(f["event"]) * 1e-3
        start_s = f["t0"].elapsed_time(f["start"]) * 1e-3
        k = len(self.batches)
        self.batch_log.append((len(f["reqs"]), start_s, done_s))
        for i, r in enumerate(f["reqs"]):
            if f["events"] is not None:
                done_s = f["t0"].elapsed_time(f["events"][i]) * 1e-3
            r.latency = done_s - (r.arrival if self._realtime else 0.0)
            r.service = done_s - start_s
            r.batch = k
            if r.keep_image:
                r.image = f["images"][i]
        for res in f["residents"]:
            self.store.release(res)
        self.batches.append(len(f["reqs"]))

    def run(self, requests: typing.List[Request], realtime: bool = True):
        """Serves `requests` (sorted by arrival). realtime: arrivals are released at their wall-clock
        times from the start of the run; otherwise all are present at t = 0."""
        self._realtime = realtime  # else every request is present at t = 0: latency from there
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue = collections.deque()
        inflight = []
        torch.cuda.synchronize()
        t0_event = torch.cuda.Event(enable_timing=True)
        t0_event.record()
        t0 = time.perf_counter()
        while pending or queue or inflight:
            now = time.perf_counter() - t0
            while pending and (not realtime or pending[0].arrival <= now):
                queue.append(pending.popleft())
            for f in [f for f in inflight if f["event"].query()]:
                inflight.remove(f)
                self._finish(f)
            if queue and len(inflight) < self.max_inflight:
                n = min(len(queue), self.max_batch)
                # small launches keep to the first stream, in arrival order (as the plain loop runs
                # them), so they do not share the device with each other; batches take the least
                # busy stream
                stream = self.streams[0] if n <= self.plain_max else self._pick_stream(inflight)
                rec, rest = self._launch([queue.popleft() for _ in range(n)], t0_event, stream)
                queue.extendleft(reversed(rest))
                if rec is not None:
                    inflight.append(rec)
                    continue
                if not inflight:  # a single scene larger than the whole budget
                    raise MemoryError(f"scene {rest[0].scene} does not fit the budget")
                oldest = inflight[0]  # launch order
                oldest["event"].synchronize()  # an in-flight batch unpins its scenes
                inflight.remove(oldest)
                self._finish(oldest)
                continue
            if inflight and (len(inflight) >= self.max_inflight or not pending):
                # no slot, or nothing left to admit: wait briefly for a batch to complete
                if not any(f["event"].query() for f in inflight):
                    time.sleep(2e-5)
                continue
            if pending:  # a slot is free: wait for the next arrival
                dt = pending[0].arrival - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(min(dt, 2e-4 if inflight else 1e-3))
        torch.cuda.synchronize()
        return requests


class PlainServer:
    """The per-request baseline: each request is rendered by its own rt_render launch as it
    arrives, in arrival order on one stream (scenes resident in the same store)."""

    def __init__(self, store: SceneStore, stream=None):
        self.store = store
        self.stream = stream or torch.cuda.Stream()

    def run(self, requests: typing.List[Request], realtime: bool = True):
        pending = sorted(requests, key=lambda r: r.arrival)
        torch.cuda.synchronize()
        t0_event = torch.cuda.Event(enable_timing=True)
        t0_event.record()
        t0 = time.perf_counter()
        done = []
        for r in pending:
            if realtime:
                dt = r.arrival - (time.perf_counter() - t0)
                while dt > 0:
                    time.sleep(min(dt, 2e-4))
                    dt = r.arrival - (time.perf_counter() - t0)
            res = self.store.acquire(r.scene)
            self.stream.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(self.stream):
                start = torch.cuda.Event(enable_timing=True)
                start.record(self.stream)
                img = res.raycaster.render(camera_for(r.eye), r.size, r.size, textures=res.binding)
                ev = torch.cuda.Event(enable_timing=True)
                ev.record(self.stream)
            self.store.release(res)
            done.append((r, start, ev, img if r.keep_image else None))
        torch.cuda.synchronize()
        for r, start, ev, img in done:
            done_s = t0_event.elapsed_time(ev) * 1e-3
            r.latency = done_s - (r.arrival if realtime else 0.0)
            r.service = done_s - t0_event.elapsed_time(start) * 1e-3
            if r.keep_image:
                r.image = img
        return requests


def poisson_arrivals(rate: float, n: int, seed: int) -> np.ndarray:
    """Arrival times (s) of a Poisson process of `rate` requests/s: cumulative exponential gaps."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


def make_requests(n: int, n_scenes: int, rate: float, seed: int, sizes=(256, 512, 1024)):
    """The serving workload: scene uniform over the pool, eye = 3 d with d uniform on the unit
    sphere (looking at the origin), image side uniform over `sizes`, Poisson arrivals."""
    rng = np.random.default_rng(seed)
    arr = poisson_arrivals(rate, n, seed + 1)
    out = []
    for i in range(n):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        out.append(Request(int(rng.integers(n_scenes)), tuple(3.0 * d), int(rng.choice(sizes)),
                           float(arr[i])))
    return out


def latency_summary(reqs, realtime=True):
    lat = np.array([r.latency for r in reqs]) * 1e3
    svc = np.array([r.service for r in reqs]) * 1e3
    if realtime:
        span = max(r.arrival + r.latency for r in reqs) - min(r.arrival for r in reqs)
    else:  # every request present at t = 0: the makespan
        span = max(r.latency for r in reqs)
    return {"requests": len(reqs), "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "max_ms": round(float(lat.max()), 3),
            "batch_gpu_ms_p50": round(float(np.percentile(svc, 50)), 3),
            "batch_gpu_ms_p99": round(float(np.percentile(svc, 99)), 3),
            "throughput_rps": round(len(reqs) / span, 1)}
