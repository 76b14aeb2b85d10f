"""Batched rendering: many independent (scene, camera, image) jobs in one persistent launch.

The reference renders one image per call (its kernels are dispatched one at a time through
rendering/_core.py:247-299). Here a batch is a device array of job records (rt_render_job: the
arguments of one rt_render call) rendered by rt_render_batch: a prepare kernel and one persistent
kernel whose waves pull 8x8 tiles of any job from a global queue, so that the long dependent
chains of the hardest tiles of all images overlap instead of running one frame after another.
Output is bit-identical to rendering each job with Raycaster.render.

  reqs = [RenderRequest(rc_i, camera_i, w_i, h_i) for ...]
  b = BatchRender(reqs)             # device job table, outputs, workspace
  b.launch()                        # or: b.capture(); b.replay()  (one hipGraph launch)
  b.images[i]                       # [h_i, w_i, 3] float32
  cfg, table = autotune(b)          # waves per CU x tiles per grab x traversals
"""
import ctypes
import dataclasses
import time
import typing

import numpy as np
import torch

from . import _native as N


@dataclasses.dataclass
class RenderRequest:
    """Arguments of one Raycaster.render call."""
    raycaster: typing.Any
    camera: typing.Any
    width: int
    height: int
    light: typing.Any = None
    textures: typing.Any = None
    mesh_textures: typing.Any = None
    shadows: bool = True
    bilinear: bool = True
    ambient: float = 0.2
    background: typing.Any = (0.0, 0.0, 0.0)
    shadow_tmin: float = 1e-4
    out: typing.Any = None
    return_hits: bool = False


_ALG = {"packet": N.RT_TRAVERSAL_PACKET, "lane": N.RT_TRAVERSAL_LANE, "hybrid": N.RT_TRAVERSAL_HYBRID,
        "auto": N.RT_TRAVERSAL_AUTO, "quad": N.RT_TRAVERSAL_QUAD}


@dataclasses.dataclass
class BatchConfig:
    """Launch shape of rt_render_batch (rt_batch_config)."""
    waves_per_cu: int = 16
    tiles_per_grab: int = 4
    shadow: str = "lane"
    primary: str = "lane"
    lane_stack: int = N.RT_LANE_STACK_MAX
    packet_budget: int = 256

    def to_c(self):
        c = N.rt_batch_config()
        c.waves_per_cu = int(self.waves_per_cu)
        c.tiles_per_grab = int(self.tiles_per_grab)
        c.primary = _ALG[self.primary]
        c.shadow = _ALG[self.shadow]
        c.lane_stack = int(self.lane_stack)
        c.packet_budget = int(self.packet_budget)
        return c

    def key(self):
        return f"w{self.waves_per_cu}g{self.tiles_per_grab}{self.primary[0]}{self.shadow[0]}"


class BatchRender:
    """A batch of render jobs resident on the device (job table, outputs, workspace).

    `images[i]` / `hits[i]` / `stats[i]` are the outputs of request i. Relaunching (launch() or a
    captured graph's replay()) re-renders every job with the current contents of the inputs."""

    def __init__(self, requests: typing.List[RenderRequest], config: BatchConfig = None,
                 stats: bool = True, stream=None):
        lib = N.require_gpu()
        if not requests:
            raise ValueError("BatchRender needs at least one request")
        if len(requests) > N.RT_BATCH_MAX_JOBS:
            raise ValueError(f"at most {N.RT_BATCH_MAX_JOBS} jobs per batch")
        self.config = config or BatchConfig()
        self._stream = stream
        dev = torch.device("cuda")
        n = len(requests)
        ts = self._torch_stream()
        cur = torch.cuda.current_stream()
        if ts != cur:
            ts.wait_stream(cur)  # inputs queued on the caller's stream are ready
        # every buffer and copy below is made on the stream the launches run on
        with torch.cuda.stream(ts):
            self.stats = torch.zeros((n, 4), dtype=torch.int32, device=dev) if stats else None
            jobs = (N.rt_render_job * n)()
            self.images, self.hits, self._keep = [], [], []
            self.rays_primary = 0
            for i, r in enumerate(requests):
                if r.width * r.height > N.RT_BATCH_MAX_JOB_PIXELS:
                    raise ValueError(f"request {i}: more than {N.RT_BATCH_MAX_JOB_PIXELS} pixels")
                p, o, out, hits, keep = r.raycaster._render_args(
                    r.width, r.height, r.light, r.textures, r.mesh_textures, r.shadows, r.bilinear,
                    r.ambient, r.background, r.shadow_tmin, r.out, r.return_hits,
                    self.stats[i] if stats else None)
                jobs[i].scene = r.raycaster._scene
                jobs[i].cam = r.camera.to_c()
                jobs[i].params = p
                jobs[i].out = o
                self.images.append(out)
                self.hits.append(hits)
                self._keep += keep + [r.raycaster]
                self.rays_primary += r.width * r.height
            self.n_jobs = n
            # pinned staging + asynchronous copy: a pageable copy would wait for every batch already
            # queued on the stream, and a serving loop enqueues the next batch behind a running one
            self._jobs_host = torch.frombuffer(bytearray(bytes(jobs)), dtype=torch.uint8).pin_memory()
            self.jobs_dev = self._jobs_host.to(dev, non_blocking=True)
            wsb = lib.rt_render_batch_workspace_bytes(n)
            self.workspace = torch.zeros(wsb, dtype=torch.uint8, device=dev)
        self._graph = None

    def _torch_stream(self):
        return self._stream if self._stream is not None else torch.cuda.current_stream()

    @property
    def stream(self):
        """The stream the batch's launches run on (its outputs are ready there)."""
        return self._torch_stream()

    def launch(self, config: BatchConfig = None):
        """Renders every job (asynchronous on the stream)."""
        lib = N.load()
        cfg = (config or self.config).to_c()
        N.check(lib.rt_render_batch(self.jobs_dev.data_ptr(), self.n_jobs, ctypes.byref(cfg),
                                    self.workspace.data_ptr(), self.workspace.numel(),
                                    N.stream_handle(self._stream)), "rt_render_batch")

    def capture(self):
        """Captures one launch (prepare + persistent kernel) into a hipGraph; replay() runs it."""
        self.launch()  # warm up outside the capture
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=self._stream):
            self.launch()
        self._graph = g
        return g

    def replay(self):
        if self._graph is None:
            raise RuntimeError("capture() first")
        self._graph.replay()

    def status(self) -> int:
        """0 if every job was rendered; bit 1: a job exceeded RT_BATCH_MAX_JOB_PIXELS, bit 2: a job
        lacked a required buffer (synchronises)."""
        return int(self.workspace[4:8].view(torch.int32).item())


class BatchSlot:
    """Reusable buffers for launching batches back to back on one stream without per-batch
    allocation of the job table: a pinned host table, its device copy and the batch workspace,
    sized for `max_jobs`. `launch(requests)` fills the tables from the requests' render
    arguments, copies them asynchronously and launches rt_render_batch; the output images are
    allocated per request (torch's caching allocator) and returned. Before the host table is
    refilled, the previous launch's copy out of it is waited for (an event after the copy)."""

    def __init__(self, max_jobs: int, stream=None, config: BatchConfig = None):
        lib = N.require_gpu()
        self.max_jobs = int(max_jobs)
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        self.config = config or BatchConfig()
        nbytes = ctypes.sizeof(N.rt_render_job) * self.max_jobs
        self._jobs = (N.rt_render_job * self.max_jobs)()
        self.host = torch.empty(nbytes, dtype=torch.uint8).pin_memory()
        with torch.cuda.stream(self.stream):
            self.dev = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
            self.workspace = torch.empty(lib.rt_render_batch_workspace_bytes(self.max_jobs),
                                         dtype=torch.uint8, device="cuda")
        self._copied = None
        self._keep = []

    def launch(self, requests: typing.List[RenderRequest], config: BatchConfig = None):
        """Renders the requests (at most max_jobs) in one launch on the slot's stream; returns
        their images [h, w, 3] float32 (ready when the stream reaches the launch's end)."""
        lib = N.load()
        n = len(requests)
        if not 0 < n <= self.max_jobs:
            raise ValueError(f"1..{self.max_jobs} requests per launch")
        if self._copied is not None:
            self._copied.synchronize()  # the previous table left the host buffer
        images, keep = [], []
        with torch.cuda.stream(self.stream):
            for i, r in enumerate(requests):
                if r.width * r.height > N.RT_BATCH_MAX_JOB_PIXELS:
                    raise ValueError(f"request {i}: more than {N.RT_BATCH_MAX_JOB_PIXELS} pixels")
                p, o, out, hits, k = r.raycaster._render_args(
                    r.width, r.height, r.light, r.textures, r.mesh_textures, r.shadows, r.bilinear,
                    r.ambient, r.background, r.shadow_tmin, r.out, r.return_hits, None)
                j = self._jobs[i]
                j.scene = r.raycaster._scene
                j.cam = r.camera.to_c()
                j.params = p
                j.out = o
                images.append(out)
                keep += k
            nbytes = ctypes.sizeof(N.rt_render_job) * n
            ctypes.memmove(self.host.data_ptr(), ctypes.addressof(self._jobs), nbytes)
            self.dev[:nbytes].copy_(self.host[:nbytes], non_blocking=True)
            self._copied = torch.cuda.Event()
            self._copied.record(self.stream)
            cfg = (config or self.config).to_c()
            N.check(lib.rt_render_batch(self.dev.data_ptr(), n, ctypes.byref(cfg), self.workspace.data_ptr(),
                                        self.workspace.numel(), N.stream_handle(self.stream)),
                    "rt_render_batch")
        self._keep = keep  # alive until the next launch on this slot's stream
        return images


def time_launches(fn, reps, warmup=1):
    """Per-call durations (ms) of `fn` from HIP events on the current stream."""
    for _ in range(warmup):
        fn()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    torch.cuda.synchronize()
    ev[0].record()
    for k in range(reps):
        fn()
        ev[k + 1].record()
    torch.cuda.synchronize()
    return [ev[k].elapsed_time(ev[k + 1]) for k in range(reps)]


DEFAULT_CANDIDATES = ([BatchConfig(w, g, "lane", "quad") for w in (8, 16, 32) for g in (1, 4, 16)] +
                      [BatchConfig(w, g, s, p) for p in ("auto", "packet") for s in ("lane", "packet")
                       for w in (16, 32) for g in (1, 4)] +
                      [BatchConfig(w, g, "lane", "lane") for w in (16, 32) for g in (1, 4, 16)])


def autotune(batch: BatchRender, candidates=None, reps=3):
    """Times every candidate launch shape on this batch (median of `reps` launches after one
    warm-up) and sets batch.config to the fastest. Returns (best, [(config, median_ms), ...])."""
    table = []
    for c in candidates or DEFAULT_CANDIDATES:
        ms = time_launches(lambda: batch.launch(c), reps)
        table.append((c, float(np.median(ms))))
    best = min(table, key=lambda x: x[1])[0]
    batch.config = best
    return best, table
