"""The C ABI boundary: librtx_hip.so loads (no GPU needed to load it) and exports every function
that include/rendertoy_rt.h declares; the ctypes mirrors in rendertoy_amd/_native.py have the
sizes and field offsets the C compiler gives the header's structs."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rendertoy_rt.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(rt_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for n in ("rt_abi_version", "rt_last_error", "rt_scene_sizes_for", "rt_scene_build",
              "rt_trace_closest", "rt_trace_any", "rt_trace_camera", "rt_generate_rays",
              "rt_render", "rt_render_workspace_bytes"):
        assert n in names


def test_library_loads_and_exports_every_declared_symbol():
    from rendertoy_amd import _native
    lib = _native.load()
    for name in declared_functions():
        assert hasattr(lib, name), f"librtx_hip.so does not export {name}"
    assert lib.rt_abi_version() == _native.RT_ABI_VERSION
    # every declared function has a ctypes signature
    sig = {n for n, _, _ in _native.SIGNATURES}
    assert set(declared_functions()) <= sig


def test_sizes_query_and_error_path_without_gpu():
    from rendertoy_amd import _native
    lib = _native.load()
    s = _native.rt_scene_sizes()
    assert lib.rt_scene_sizes_for(1000, ctypes.byref(s)) == 0
    assert s.nodes_bytes == 64 * 999 and s.nodes4_bytes == 128 * 999
    assert s.tris_bytes == 48 * 1000 and s.shade_bytes == 64 * 1000
    assert s.workspace_bytes > 0
    # too many triangles: RT_ECAPACITY and a message, no device work
    assert lib.rt_scene_sizes_for(1 << 30, ctypes.byref(s)) == -3
    assert b"too many" in lib.rt_last_error()
    # NULL scene pointer is rejected before any launch
    assert lib.rt_trace_closest(None, None, 1, None, None, None) == -1


PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "rendertoy_rt.h"
#define S(T) printf(#T " %zu\n", sizeof(T));
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
    S(rt_mesh_desc) O(rt_mesh_desc, vertices) O(rt_mesh_desc, indices) O(rt_mesh_desc, n_vertices)
    O(rt_mesh_desc, n_tris) O(rt_mesh_desc, vertex_stride) O(rt_mesh_desc, pos_offset)
    O(rt_mesh_desc, nrm_offset) O(rt_mesh_desc, uv_offset)
    S(rt_scene) O(rt_scene, root) O(rt_scene, nodes) O(rt_scene, nodes4) O(rt_scene, tris)
    O(rt_scene, shade) O(rt_scene, mesh_offsets) O(rt_scene, bounds) O(rt_scene, prim_slot)
    S(rt_scene_sizes) O(rt_scene_sizes, nodes4_bytes) O(rt_scene_sizes, workspace_bytes)
    O(rt_scene_sizes, prim_slot_bytes)
    S(rt_rays) O(rt_rays, tmin) O(rt_rays, stride) O(rt_rays, tmin_default) O(rt_rays, tmax_default)
    S(rt_hits) O(rt_hits, mesh) O(rt_hits, stride)
    S(rt_camera) O(rt_camera, up) O(rt_camera, forward)
    S(rt_texture) O(rt_texture, width)
    S(rt_render_params) O(rt_render_params, light) O(rt_render_params, ambient)
    O(rt_render_params, shadow_tmin) O(rt_render_params, textures_dev)
    O(rt_render_params, mesh_texture_dev) O(rt_render_params, n_textures)
    S(rt_render_out) O(rt_render_out, stats)
    S(rt_tuning) O(rt_tuning, shadow) O(rt_tuning, super_tile) O(rt_tuning, lane_stack)
    S(rt_render_job) O(rt_render_job, cam) O(rt_render_job, params) O(rt_render_job, out)
    S(rt_batch_config) O(rt_batch_config, tiles_per_grab) O(rt_batch_config, primary)
    O(rt_batch_config, shadow) O(rt_batch_config, lane_stack) O(rt_batch_config, packet_budget)
    O(rt_tuning, packet_budget) O(rt_tuning, builder) O(rt_tuning, query)
    O(rt_tuning, handoff_steps) O(rt_tuning, item_steps)
    S(rt_paths_params) O(rt_paths_params, spp) O(rt_paths_params, seed) O(rt_paths_params, albedo)
    O(rt_paths_params, ray_tmin) S(rt_paths_out) O(rt_paths_out, counts)
    return 0;
}
"""


def test_ctypes_layouts_match_the_c_header(tmp_path):
    from rendertoy_amd import _native as N
    src = tmp_path / "probe.c"
    src.write_text(PROBE)
    exe = tmp_path / "probe"
    r = subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    c = {}
    for line in out.splitlines():
        k, v = line.split()
        c[k] = int(v)
    for name in ("rt_mesh_desc", "rt_scene", "rt_scene_sizes", "rt_rays", "rt_hits", "rt_camera",
                 "rt_texture", "rt_render_params", "rt_render_out", "rt_tuning", "rt_render_job",
                 "rt_batch_config", "rt_paths_params", "rt_paths_out"):
        cls = getattr(N, name)
        assert ctypes.sizeof(cls) == c[name], name
        for key, off in c.items():
            if key.startswith(name + "."):
                field = key.split(".", 1)[1]
                assert getattr(cls, field).offset == off, key


def test_tuning_get_set_and_range_checks_without_gpu():
    import rendertoy_amd as ren
    from rendertoy_amd import _native
    d = ren.get_tuning()
    assert d == {"primary": _native.RT_TRAVERSAL_QUAD, "shadow": _native.RT_TRAVERSAL_LANE,
                 "super_tile": 0, "lane_stack": _native.RT_LANE_STACK_MAX, "packet_budget": 256,
                 "builder": _native.RT_BUILD_LBVH, "query": _native.RT_TRAVERSAL_LANE,
                 "handoff_steps": 0, "item_steps": 32}
    with ren.tuning(primary="hybrid", lane_stack=3, super_tile=2, packet_budget=7, builder="ploc",
                    query="binary", item_steps=5):
        assert ren.get_tuning() == {"primary": 2, "shadow": 1, "super_tile": 2, "lane_stack": 3,
                                    "packet_budget": 7, "builder": _native.RT_BUILD_PLOC,
                                    "query": _native.RT_TRAVERSAL_BINARY, "handoff_steps": 0,
                                    "item_steps": 5}
    assert ren.get_tuning() == d
    for bad in ({"lane_stack": 0}, {"lane_stack": 33}, {"primary": 5}, {"shadow": 3},
                {"super_tile": 65}, {"packet_budget": 0}, {"builder": 2}, {"query": "packet"},
                {"query": "quad"}, {"item_steps": 0}, {"handoff_steps": 1}):
        with pytest.raises(RuntimeError, match="rt_set_tuning"):
            ren.set_tuning(**bad)
    assert ren.get_tuning() == d
    with pytest.raises(TypeError):
        ren.set_tuning(tile=4)


def test_per_call_tuning_is_checked_before_any_launch_without_gpu():
    """A per-call rt_tuning is range-checked like rt_set_tuning, before any device work (the
    fake addresses below are never touched), and NULL means the process default."""
    from rendertoy_amd import _native as N
    from rendertoy_amd.tuning import make_tuning
    lib = N.load()
    sc = N.rt_scene()
    sc.root = N.RT_EMPTY_ROOT
    cam = N.rt_camera()
    hits = N.rt_hits(0x1000, None, None, None, None, None, 4)
    rays = N.rt_rays(0x1000, 0x1000, 0x1000, 0x1000, 0x1000, 0x1000, None, None, 4, 0.0, 1.0)
    for bad in (dict(lane_stack=0), dict(primary=7), dict(query=N.RT_TRAVERSAL_PACKET), dict(builder=3)):
        t = make_tuning(None, **bad)
        assert lib.rt_trace_camera(ctypes.byref(sc), ctypes.byref(cam), 8, 8, ctypes.byref(hits),
                                   None, 0, ctypes.byref(t), None) == -1
        assert b"rt_trace_camera: tuning" in lib.rt_last_error()
        assert lib.rt_trace_closest(ctypes.byref(sc), ctypes.byref(rays), 8, ctypes.byref(hits),
                                    ctypes.byref(t), None) == -1
        assert lib.rt_trace_any(ctypes.byref(sc), ctypes.byref(rays), 8, 0x1000, ctypes.byref(t), None) == -1
        assert b"rt_trace_any: tuning" in lib.rt_last_error()
    # nothing to trace: no launch, whatever the tuning
    assert lib.rt_trace_camera(ctypes.byref(sc), ctypes.byref(cam), 0, 8, ctypes.byref(hits), None, 0,
                               None, None) == 0


def test_export_and_texel_entries_reject_bad_arguments_without_gpu():
    from rendertoy_amd import _native as N
    lib = N.load()
    sc = N.rt_scene()
    sc.n_tris, sc.n_meshes, sc.root = 10, 1, 0
    sc.nodes, sc.bounds, sc.prim_slot = 0x1000, 0x2000, 0x6000
    m = N.rt_mesh_desc(0x3000, None, 30, 10, 80, 0, N.RT_ABSENT, N.RT_ABSENT)
    assert lib.rt_scene_export(None, None, 0, 0x4000, 0x5000, None, None) == -1
    assert lib.rt_scene_export(ctypes.byref(sc), ctypes.byref(m), 1, 0x4008, 0x5000, None, None) == -1
    assert b"16-byte" in lib.rt_last_error()
    assert lib.rt_scene_export(ctypes.byref(sc), ctypes.byref(m), 2, 0x4000, 0x5000, None, None) == -1
    m.n_tris = 9
    assert lib.rt_scene_export(ctypes.byref(sc), ctypes.byref(m), 1, 0x4000, 0x5000, None, None) == -1
    assert b"do not match" in lib.rt_last_error()
    m.n_tris = 10
    sc.prim_slot = None
    assert lib.rt_scene_export(ctypes.byref(sc), ctypes.byref(m), 1, 0x4000, 0x5000, None, None) == -1
    assert lib.rt_texels_to_rgba8(None, 0, None, None) == 0
    assert lib.rt_texels_to_rgba8(None, 4, 0x1000, None) == -1
    assert lib.rt_texels_to_rgba8(0x1008, 4, 0x2000, None) == -1
    assert lib.rt_texels_to_rgba8(0x1000, 4, 0x2002, None) == -1


def test_batch_entry_rejects_bad_arguments_without_gpu():
    from rendertoy_amd import _native as N
    lib = N.load()
    cfg = N.rt_batch_config(8, 1, N.RT_TRAVERSAL_HYBRID, N.RT_TRAVERSAL_LANE, 32, 256)
    ws = ctypes.create_string_buffer(1 << 16)
    wsp = ctypes.cast(ws, ctypes.c_void_p).value
    fake_jobs = 0x1000
    assert lib.rt_render_batch(None, 1, ctypes.byref(cfg), wsp, 1 << 16, None) == -1
    assert lib.rt_render_batch(fake_jobs, 1, None, wsp, 1 << 16, None) == -1
    assert lib.rt_render_batch(fake_jobs, N.RT_BATCH_MAX_JOBS + 1, ctypes.byref(cfg), wsp, 1 << 20, None) == -3
    for bad in (N.rt_batch_config(0, 1, 0, 1, 32, 256), N.rt_batch_config(33, 1, 0, 1, 32, 256),
                N.rt_batch_config(8, 0, 0, 1, 32, 256), N.rt_batch_config(8, 65, 0, 1, 32, 256),
                N.rt_batch_config(8, 1, 5, 1, 32, 256), N.rt_batch_config(8, 1, 0, 3, 32, 256),
                N.rt_batch_config(8, 1, 0, 1, 0, 256), N.rt_batch_config(8, 1, 0, 1, 32, 0)):
        assert lib.rt_render_batch(fake_jobs, 1, ctypes.byref(bad), wsp, 1 << 16, None) == -1
    need = lib.rt_render_batch_workspace_bytes(10)
    assert need == 4 * (64 + 11)
    assert lib.rt_render_batch(fake_jobs, 10, ctypes.byref(cfg), wsp, need - 1, None) == -1
    assert b"workspace" in lib.rt_last_error()
    assert lib.rt_render_batch(fake_jobs, 0, ctypes.byref(cfg), None, 0, None) == 0  # nothing to do


def test_paths_entry_rejects_bad_arguments_without_gpu():
    from rendertoy_amd import _native as N
    lib = N.load()
    sc = N.rt_scene()
    sc.root = N.RT_EMPTY_ROOT
    cam = N.rt_camera()
    out = N.rt_paths_out(0x1000, None)
    ws = 0x10000
    ok = N.rt_paths_params(8, 8, 4, 2, 1, 0, 0.7, 1.0, 1e-4)
    need = lib.rt_paths_workspace_bytes(8, 8, 4)
    assert need == 256 + 4 * 256 + 2 * 48 * 256  # counters, samples, two queues
    assert lib.rt_trace_paths(None, ctypes.byref(cam), ctypes.byref(ok), ctypes.byref(out), ws, need, None, None) == -1
    for bad in (N.rt_paths_params(8, 8, 0, 2, 1, 0, 0.7, 1.0, 1e-4),
                N.rt_paths_params(8, 8, 4, 17, 1, 0, 0.7, 1.0, 1e-4)):
        assert lib.rt_trace_paths(ctypes.byref(sc), ctypes.byref(cam), ctypes.byref(bad),
                                  ctypes.byref(out), ws, need, None, None) == -1
    assert lib.rt_trace_paths(ctypes.byref(sc), ctypes.byref(cam), ctypes.byref(ok), ctypes.byref(out),
                              ws, need - 1, None, None) == -1
    assert lib.rt_trace_paths(ctypes.byref(sc), ctypes.byref(cam), ctypes.byref(ok), ctypes.byref(out),
                              ws + 16, need, None, None) == -1
    assert b"256-byte" in lib.rt_last_error()
    big = N.rt_paths_params(65536, 65536, 1, 2, 1, 0, 0.7, 1.0, 1e-4)
    assert lib.rt_trace_paths(ctypes.byref(sc), ctypes.byref(cam), ctypes.byref(big), ctypes.byref(out),
                              ws, need, None, None) == -3


def test_mesh_and_image_entries_reject_bad_arguments_without_gpu():
    """Argument checks of rt_mesh_compute_normals / rt_mesh_pack / rt_mesh_unpack /
    rt_image_to_bgra8: each case returns RT_EINVAL (or RT_OK for empty work) before any device work,
    so the fake device addresses below are never touched."""
    from rendertoy_amd import _native as N
    lib = N.load()
    EINVAL = -1
    dev, out = 0x10000, 0x20000

    def mesh(nv, nt, vertices=dev, indices=None, stride=80, pos=0):
        return N.rt_mesh_desc(vertices, indices, nv, nt, stride, pos, N.RT_ABSENT, N.RT_ABSENT)

    def normals(m, o=out, stride=12, ws=None, ws_bytes=0):
        return lib.rt_mesh_compute_normals(None if m is None else ctypes.byref(m), o, stride, ws, ws_bytes, None)

    assert normals(None) == EINVAL
    assert normals(mesh(0, 0)) == 0                      # nothing to do
    assert normals(mesh(0, 1)) == EINVAL                 # triangles without vertices
    assert normals(mesh(3, 1, vertices=None)) == EINVAL
    assert normals(mesh(3, 1, stride=8)) == EINVAL       # stride below one position
    assert normals(mesh(3, 1, stride=16, pos=8)) == EINVAL
    assert normals(mesh(3, 2)) == EINVAL                 # non-indexed: 3 vertices per triangle
    assert normals(mesh(3, 1), o=None) == EINVAL
    assert normals(mesh(3, 1), stride=8) == EINVAL
    assert normals(mesh(3, 1), o=out + 2) == EINVAL      # misaligned output
    need = lib.rt_mesh_normals_workspace_bytes(4, 2)
    assert need > 0
    assert normals(mesh(4, 2, indices=0x30000), ws=0x40000, ws_bytes=need - 1) == EINVAL
    assert b"workspace too small" in lib.rt_last_error()
    assert normals(mesh(4, 2, indices=0x30000), ws=0x40010, ws_bytes=need) == EINVAL
    assert b"256-byte" in lib.rt_last_error()

    assert lib.rt_mesh_pack(None, None, None, 0, None, None) == 0
    assert lib.rt_mesh_pack(None, None, None, 4, dev, None) == EINVAL
    assert lib.rt_mesh_pack(dev, dev + 1, None, 4, out, None) == EINVAL
    assert lib.rt_mesh_pack(dev, None, None, 4, out + 8, None) == EINVAL
    assert lib.rt_mesh_unpack(None, 0, None, None, None, None) == 0
    assert lib.rt_mesh_unpack(None, 4, dev, None, None, None) == EINVAL
    assert lib.rt_mesh_unpack(out + 8, 4, dev, None, None, None) == EINVAL
    assert lib.rt_mesh_unpack(out, 4, dev + 2, None, None, None) == EINVAL

    assert lib.rt_image_to_bgra8(None, 0, None, None) == 0
    assert lib.rt_image_to_bgra8(None, 4, out, None) == EINVAL
    assert lib.rt_image_to_bgra8(dev, 4, out + 1, None) == EINVAL
