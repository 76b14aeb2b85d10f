"""Launch choices of the build, trace and render kernels: which traversal walks the primary, the
shadow and the ray-buffer rays (packet, per-lane, or packet then per-lane after a step budget),
the block -> XCD mapping, the per-lane stack size, the hybrid's packet budget and the hierarchy
builder. Every choice returns the same hits; they trade speed only. The reference has no
counterpart (its kernels run a fixed 32-work-item group, rendering/_core.py:286).

Each entry point takes its choices per call: `Raycaster(models, tuning={...})` fixes them for one
scene (its build and every query and render of it), and the query/render methods also take a
`tuning=` of their own. A call without one uses the process default, which `set_tuning` changes
(and `tuning()` changes for the duration of a block); the default is copied at the call, so a
caller that passes its own choices never races with another that changes the default."""
import contextlib
import ctypes

from . import _native

PACKET = _native.RT_TRAVERSAL_PACKET
LANE = _native.RT_TRAVERSAL_LANE
HYBRID = _native.RT_TRAVERSAL_HYBRID
AUTO = _native.RT_TRAVERSAL_AUTO
QUAD = _native.RT_TRAVERSAL_QUAD
BINARY = _native.RT_TRAVERSAL_BINARY
_NAMES = {"packet": PACKET, "lane": LANE, "hybrid": HYBRID, "auto": AUTO, "quad": QUAD,
          "binary": BINARY}
LBVH = _native.RT_BUILD_LBVH
PLOC = _native.RT_BUILD_PLOC
_BUILDERS = {"lbvh": LBVH, "ploc": PLOC}
FIELDS = ("primary", "shadow", "super_tile", "lane_stack", "packet_budget", "builder", "query",
          "handoff_steps", "item_steps")


def get_tuning():
    """The process default as a dict {primary, shadow, super_tile, lane_stack, packet_budget,
    builder, query, handoff_steps, item_steps}."""
    lib = _native.load()
    t = _native.rt_tuning()
    _native.check(lib.rt_get_tuning(t), "rt_get_tuning")
    return {k: getattr(t, k) for k in FIELDS}


def _merged(base, kw):
    cur = dict(base)
    for k, v in kw.items():
        if k not in cur:
            raise TypeError(f"unknown tuning field {k!r}")
        if k in ("primary", "shadow", "query"):
            v = _NAMES.get(v, v)
        elif k == "builder":
            v = _BUILDERS.get(v, v)
        cur[k] = int(v)
    return cur


def make_tuning(base=None, **kw):
    """A complete rt_tuning: `base` (a dict, default: the process default) with the fields of `kw`
    replaced. Traversals may be given as "packet"/"lane"/"hybrid"/"auto"/"quad"/"binary" or the
    constants, the builder as "lbvh"/"ploc". Values are range-checked by the library when the
    tuning is used (RuntimeError with its message)."""
    cur = _merged(base if base is not None else get_tuning(), kw)
    return _native.rt_tuning(*[cur[k] for k in FIELDS])


def as_dict(t):
    return {k: getattr(t, k) for k in FIELDS}


def tuning_arg(t):
    """The `const rt_tuning*` argument of an entry point: NULL (process default) for None."""
    return None if t is None else ctypes.byref(t)


def set_tuning(**kw):
    """Updates fields of the process default (see make_tuning for the accepted values).
    Raises RuntimeError (the library's message) for out-of-range values."""
    lib = _native.load()
    t = make_tuning(None, **kw)
    _native.check(lib.rt_set_tuning(t), "rt_set_tuning")


@contextlib.contextmanager
def tuning(**kw):
    """Context manager: set_tuning(**kw) inside the block, previous default restored after."""
    old = get_tuning()
    set_tuning(**kw)
    try:
        yield
    finally:
        set_tuning(**old)
