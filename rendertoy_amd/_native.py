"""ctypes bindings of librtx_hip.so (declarations: include/rendertoy_rt.h).

The library is loaded after `import torch` so that its HIP runtime dependency resolves to the one
torch already loaded (one runtime per process). There is no fallback: if the library is missing
or fails to load, every ray-casting entry point raises.
"""
import ctypes
import os

import torch  # noqa: F401  (must be imported before the HIP library; see module docstring)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librtx_hip.so")

RT_ABI_VERSION = 10
RT_ABSENT = 0xFFFFFFFF
RT_EMPTY_ROOT = 0x7FFFFFFF
RT_RENDER_SHADOWS = 1
RT_RENDER_BILINEAR = 2
RT_TRAVERSAL_PACKET = 0
RT_TRAVERSAL_LANE = 1
RT_TRAVERSAL_HYBRID = 2
RT_TRAVERSAL_AUTO = 3
RT_TRAVERSAL_QUAD = 4
RT_TRAVERSAL_BINARY = 5
RT_LANE_STACK_MAX = 32
RT_BUILD_LBVH = 0
RT_BUILD_PLOC = 1
RT_BATCH_MAX_JOBS = 4096
RT_BATCH_MAX_JOB_PIXELS = 1 << 24
RT_PATHS_MAX_BOUNCES = 16

_c_float_p = ctypes.POINTER(ctypes.c_float)
_c_i32_p = ctypes.POINTER(ctypes.c_int32)
_c_u32_p = ctypes.POINTER(ctypes.c_uint32)


class rt_mesh_desc(ctypes.Structure):
    _fields_ = [
        ("vertices", ctypes.c_void_p),
        ("indices", ctypes.c_void_p),
        ("n_vertices", ctypes.c_uint32),
        ("n_tris", ctypes.c_uint32),
        ("vertex_stride", ctypes.c_uint32),
        ("pos_offset", ctypes.c_uint32),
        ("nrm_offset", ctypes.c_uint32),
        ("uv_offset", ctypes.c_uint32),
    ]


class rt_scene(ctypes.Structure):
    _fields_ = [
        ("n_tris", ctypes.c_uint32),
        ("n_meshes", ctypes.c_uint32),
        ("root", ctypes.c_int32),
        ("_reserved", ctypes.c_uint32),
        ("nodes", ctypes.c_void_p),
        ("nodes4", ctypes.c_void_p),
        ("tris", ctypes.c_void_p),
        ("shade", ctypes.c_void_p),
        ("mesh_offsets", ctypes.c_void_p),
        ("bounds", ctypes.c_void_p),
        ("prim_slot", ctypes.c_void_p),
    ]


class rt_scene_sizes(ctypes.Structure):
    _fields_ = [
        ("nodes_bytes", ctypes.c_uint64),
        ("nodes4_bytes", ctypes.c_uint64),
        ("tris_bytes", ctypes.c_uint64),
        ("shade_bytes", ctypes.c_uint64),
        ("bounds_bytes", ctypes.c_uint64),
        ("workspace_bytes", ctypes.c_uint64),
        ("prim_slot_bytes", ctypes.c_uint64),
    ]


class rt_rays(ctypes.Structure):
    _fields_ = [
        ("ox", ctypes.c_void_p), ("oy", ctypes.c_void_p), ("oz", ctypes.c_void_p),
        ("dx", ctypes.c_void_p), ("dy", ctypes.c_void_p), ("dz", ctypes.c_void_p),
        ("tmin", ctypes.c_void_p), ("tmax", ctypes.c_void_p),
        ("stride", ctypes.c_uint32),
        ("tmin_default", ctypes.c_float),
        ("tmax_default", ctypes.c_float),
    ]


class rt_hits(ctypes.Structure):
    _fields_ = [
        ("t", ctypes.c_void_p), ("u", ctypes.c_void_p), ("v", ctypes.c_void_p),
        ("prim", ctypes.c_void_p), ("index", ctypes.c_void_p), ("mesh", ctypes.c_void_p),
        ("stride", ctypes.c_uint32),
    ]


class rt_camera(ctypes.Structure):
    _fields_ = [
        ("eye", ctypes.c_float * 3),
        ("right", ctypes.c_float * 3),
        ("up", ctypes.c_float * 3),
        ("forward", ctypes.c_float * 3),
    ]


class rt_texture(ctypes.Structure):
    _fields_ = [
        ("texels", ctypes.c_void_p),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
    ]


class rt_render_params(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("light", ctypes.c_float * 3),
        ("ambient", ctypes.c_float),
        ("background", ctypes.c_float * 3),
        ("shadow_tmin", ctypes.c_float),
        ("textures_dev", ctypes.c_void_p),
        ("mesh_texture_dev", ctypes.c_void_p),
        ("n_textures", ctypes.c_uint32),
    ]


class rt_render_out(ctypes.Structure):
    _fields_ = [
        ("rgb", ctypes.c_void_p),
        ("t", ctypes.c_void_p),
        ("prim", ctypes.c_void_p),
        ("stats", ctypes.c_void_p),
    ]


class rt_tuning(ctypes.Structure):
    _fields_ = [
        ("primary", ctypes.c_int32),
        ("shadow", ctypes.c_int32),
        ("super_tile", ctypes.c_uint32),
        ("lane_stack", ctypes.c_uint32),
        ("packet_budget", ctypes.c_uint32),
        ("builder", ctypes.c_uint32),
        ("query", ctypes.c_int32),
        ("handoff_steps", ctypes.c_uint32),
        ("item_steps", ctypes.c_uint32),
    ]


class rt_render_job(ctypes.Structure):
    _fields_ = [
        ("scene", rt_scene),
        ("cam", rt_camera),
        ("params", rt_render_params),
        ("out", rt_render_out),
    ]


class rt_batch_config(ctypes.Structure):
    _fields_ = [
        ("waves_per_cu", ctypes.c_uint32),
        ("tiles_per_grab", ctypes.c_uint32),
        ("primary", ctypes.c_int32),
        ("shadow", ctypes.c_int32),
        ("lane_stack", ctypes.c_uint32),
        ("packet_budget", ctypes.c_uint32),
    ]


class rt_paths_params(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("spp", ctypes.c_uint32),
        ("bounces", ctypes.c_uint32),
        ("seed", ctypes.c_uint32),
        ("frame", ctypes.c_uint32),
        ("albedo", ctypes.c_float),
        ("sky", ctypes.c_float),
        ("ray_tmin", ctypes.c_float),
    ]


class rt_paths_out(ctypes.Structure):
    _fields_ = [
        ("radiance", ctypes.c_void_p),
        ("counts", ctypes.c_void_p),
    ]


# (name, restype, argtypes) of every exported entry point
SIGNATURES = [
    ("rt_abi_version", ctypes.c_int, []),
    ("rt_build_id", ctypes.c_char_p, []),
    ("rt_last_error", ctypes.c_char_p, []),
    ("rt_get_tuning", ctypes.c_int, [ctypes.POINTER(rt_tuning)]),
    ("rt_set_tuning", ctypes.c_int, [ctypes.POINTER(rt_tuning)]),
    ("rt_scene_sizes_for", ctypes.c_int, [ctypes.c_uint32, ctypes.POINTER(rt_scene_sizes)]),
    ("rt_scene_build", ctypes.c_int,
     [ctypes.POINTER(rt_mesh_desc), ctypes.c_uint32, ctypes.POINTER(rt_scene), ctypes.c_void_p,
      ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(rt_tuning), ctypes.c_void_p]),
    ("rt_scene_export", ctypes.c_int,
     [ctypes.POINTER(rt_scene), ctypes.POINTER(rt_mesh_desc), ctypes.c_uint32, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("rt_trace_closest", ctypes.c_int,
     [ctypes.POINTER(rt_scene), ctypes.POINTER(rt_rays), ctypes.c_uint32, ctypes.POINTER(rt_hits),
      ctypes.POINTER(rt_tuning), ctypes.c_void_p]),
    ("rt_trace_any", ctypes.c_int,
     [ctypes.POINTER(rt_scene), ctypes.POINTER(rt_rays), ctypes.c_uint32, ctypes.c_void_p,
      ctypes.POINTER(rt_tuning), ctypes.c_void_p]),
    ("rt_generate_rays", ctypes.c_int,
     [ctypes.POINTER(rt_camera), ctypes.c_uint32, ctypes.c_uint32] + [ctypes.c_void_p] * 6
     + [ctypes.c_void_p]),
    ("rt_trace_camera_workspace_bytes", ctypes.c_uint64, [ctypes.c_uint32, ctypes.c_uint32]),
    ("rt_trace_camera", ctypes.c_int,
     [ctypes.POINTER(rt_scene), ctypes.POINTER(rt_camera), ctypes.c_uint32, ctypes.c_uint32,
      ctypes.POINTER(rt_hits), ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(rt_tuning),
      ctypes.c_void_p]),
    ("rt_render_workspace_bytes", ctypes.c_uint64, [ctypes.c_uint32]),
    ("rt_render", ctypes.c_int,
     [ctypes.POINTER(rt_scene), ctypes.POINTER(rt_camera), ctypes.POINTER(rt_render_params),
      ctypes.POINTER(rt_render_out), ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(rt_tuning), ctypes.c_void_p]),
    ("rt_render_batch_workspace_bytes", ctypes.c_uint64, [ctypes.c_uint32]),
    ("rt_render_batch", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(rt_batch_config), ctypes.c_void_p,
      ctypes.c_uint64, ctypes.c_void_p]),
    ("rt_render_batch_status", ctypes.c_void_p, [ctypes.c_void_p]),
    ("rt_render_batch_static", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(rt_batch_config), ctypes.c_void_p]),
    ("rt_paths_workspace_bytes", ctypes.c_uint64, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    ("rt_trace_paths", ctypes.c_int,
     [ctypes.POINTER(rt_scene), ctypes.POINTER(rt_camera), ctypes.POINTER(rt_paths_params),
      ctypes.POINTER(rt_paths_out), ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(rt_tuning), ctypes.c_void_p]),
    ("rt_mesh_normals_workspace_bytes", ctypes.c_uint64, [ctypes.c_uint32, ctypes.c_uint32]),
    ("rt_mesh_compute_normals", ctypes.c_int,
     [ctypes.POINTER(rt_mesh_desc), ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
      ctypes.c_uint64, ctypes.c_void_p]),
    ("rt_mesh_pack", ctypes.c_int, [ctypes.c_void_p] * 3 + [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]),
    ("rt_mesh_unpack", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32] + [ctypes.c_void_p] * 4),
    ("rt_image_to_bgra8", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]),
    ("rt_texels_to_rgba8", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]),
]

_lib = None


def load(path=LIB_PATH):
    """Loads and type-annotates librtx_hip.so; raises RuntimeError if it is absent or stale."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m rendertoy_amd.build` (hipcc, gfx950). "
            "rendertoy_amd has no CPU fallback for ray casting.")
    lib = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    v = lib.rt_abi_version()
    if v != RT_ABI_VERSION:
        raise RuntimeError(f"{path}: ABI version {v}, expected {RT_ABI_VERSION}; rebuild it")
    from .build import source_id
    built, have = lib.rt_build_id().decode(), source_id()
    if built != have:
        raise RuntimeError(f"{path} was built from other sources (build id {built}, sources {have}): "
                           "rebuild it with `python -m rendertoy_amd.build`")
    if path == LIB_PATH:
        _lib = lib
    return lib


def check(code, what=""):
    """Raises RuntimeError with the library's message for a negative return code."""
    if code != 0:
        msg = _lib.rt_last_error().decode() if _lib is not None else ""
        raise RuntimeError(f"{what} failed ({code}): {msg}")


def require_gpu():
    """Returns the loaded library; raises unless a HIP device is present."""
    lib = load()
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device visible: rendertoy_amd ray casting runs only on the GPU")
    return lib


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)
