"""Launch choices of the camera and render kernels (process-wide): which traversal walks the
primary and the shadow rays (packet, per-lane, or packet then per-lane after a step budget), the
block -> XCD mapping, the per-lane stack size, the hybrid's packet budget and the hierarchy
builder of the next scene builds. Every choice
returns the same hits; they trade speed only. The reference has no counterpart (its kernels run a
fixed 32-work-item group, rendering/_core.py:286)."""
import contextlib

from . import _native

PACKET = _native.RT_TRAVERSAL_PACKET
LANE = _native.RT_TRAVERSAL_LANE
HYBRID = _native.RT_TRAVERSAL_HYBRID
AUTO = _native.RT_TRAVERSAL_AUTO
QUAD = _native.RT_TRAVERSAL_QUAD
_NAMES = {"packet": PACKET, "lane": LANE, "hybrid": HYBRID, "auto": AUTO, "quad": QUAD}
LBVH = _native.RT_BUILD_LBVH
PLOC = _native.RT_BUILD_PLOC
_BUILDERS = {"lbvh": LBVH, "ploc": PLOC}


def get_tuning():
    """Current settings as a dict {primary, shadow, super_tile, lane_stack, packet_budget, builder}."""
    lib = _native.load()
    t = _native.rt_tuning()
    _native.check(lib.rt_get_tuning(t), "rt_get_tuning")
    return {"primary": t.primary, "shadow": t.shadow, "super_tile": t.super_tile,
            "lane_stack": t.lane_stack, "packet_budget": t.packet_budget, "builder": t.builder}


def set_tuning(**kw):
    """Updates the given fields; traversals may be given as "packet"/"lane"/"hybrid"/"auto"/"quad"
    ("auto" and "quad" for camera rays only) or the constants, the builder as "lbvh"/"ploc" (read by
    the next scene build).
    Raises RuntimeError (the library's message) for out-of-range values."""
    lib = _native.load()
    cur = get_tuning()
    for k, v in kw.items():
        if k not in cur:
            raise TypeError(f"unknown tuning field {k!r}")
        if k in ("primary", "shadow"):
            v = _NAMES.get(v, v)
        elif k == "builder":
            v = _BUILDERS.get(v, v)
        cur[k] = v
    t = _native.rt_tuning(int(cur["primary"]), int(cur["shadow"]), int(cur["super_tile"]),
                          int(cur["lane_stack"]), int(cur["packet_budget"]), int(cur["builder"]))
    _native.check(lib.rt_set_tuning(t), "rt_set_tuning")


@contextlib.contextmanager
def tuning(**kw):
    """Context manager: set_tuning(**kw) inside the block, previous settings restored after."""
    old = get_tuning()
    set_tuning(**kw)
    try:
        yield
    finally:
        set_tuning(**old)
