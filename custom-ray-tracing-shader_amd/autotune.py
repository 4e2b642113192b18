"""Python front end of the library's launch autotuner (RtAutotune / RtBatchAutotune).

The search runs inside libCudaRayTracer.so: for a session it times device-resident frames of
the given camera and settings over BVH width {4, 8} x leaf size {2, 4} x persistent workgroups
per CU x LDS stack x traversal loop (at two path pipelines), then pipelines, workgroups, the
path-completion bounce and stack x loop again, and keeps the fastest; for a batch of views
the same without leaf size and path completion.  Winners persist in the library's JSON cache
($RT_AUTOTUNE_CACHE, default ~/.cache/rt_amd/autotune.json) keyed by scene class, and
RtInitialize / RtRender / RtBatchBuild apply them to every tuned option the caller did not set.
"""
import ctypes as c

from native_api import Batch, Session

KEYS = ("bvh_width", "build_max_leaf", "blocks_per_cu", "short_stack", "while_while", "refill", "pipelines",
        "tail_bounce")
BATCH_KEYS = ("bvh_width", "blocks_per_cu", "short_stack", "while_while", "refill", "pipelines")


def autotune(triangles, camera, settings, width, height, frames=6, fixed=None, lib=None):
    """Searches the session launch for this scene; returns ({option: value, ...,
    "mrays_per_s": best}, number of trials).  `fixed` options are set first and kept."""
    with Session(lib) as s:
        for k, v in (fixed or {}).items():
            if not s.set_option(k, v):
                raise ValueError(f"option {k}={v} rejected")
        s.initialize(triangles, width, height)
        mr, n = c.c_float(), c.c_int()
        s.check(s.lib.RtAutotune(s.handle, c.byref(camera), c.byref(settings), frames, c.byref(mr), c.byref(n)))
        best = {k: s.get_option(k) for k in KEYS}
    best["mrays_per_s"] = round(mr.value, 2)
    return best, n.value


def apply(session, cfg):
    for key in KEYS:
        if key in cfg and key not in ("bvh_width", "build_max_leaf"):
            if not session.set_option(key, cfg[key]):
                raise ValueError(f"option {key}={cfg[key]} rejected")


def autotune_batch(views, frames=4, fixed=None, lib=None):
    """Searches the batch launch for these views; returns (best options, number of trials)."""
    with Batch(lib, **(fixed or {})) as b:
        for t, w, h, cam, st in views:
            b.add_view(t, w, h, cam, st)
        mr, n = c.c_float(), c.c_int()
        b.check(b.lib.RtBatchAutotune(b.handle, frames, c.byref(mr), c.byref(n)))
        best = {k: b.lib.RtBatchGetOption(b.handle, k.encode()) for k in BATCH_KEYS}
    best["mrays_per_s"] = round(mr.value, 2)
    return best, n.value
