"""Autotuner for the traversal launch: BVH width and leaf size, persistent workgroups per
CU, LDS short-stack depth, the traversal loop (speculative while-while with per-lane ray
refill, or one step per iteration), the number of concurrent path pipelines and the bounce
from which paths complete in one launch.

For a scene, camera and settings it initialises one session per (BVH width, leaf size),
times a few device-resident frames for every (blocks_per_cu, short_stack, loop) combination,
then, on the fastest of those, tries each pipeline count, each workgroup count, each
path-completion bounce and again each (short_stack, loop), and keeps the configuration with
the highest ray throughput.  Choices persist in a JSON cache keyed by
scene size class and resolution ($RT_AUTOTUNE_CACHE, default ~/.cache/rt_amd/autotune.json),
so later sessions of a similar scene reuse them without searching again.
"""
import json
import math
import os
import time
from pathlib import Path

from native_api import Batch, Session

WIDTHS = (4, 8)
MAX_LEAF = (4,)
BLOCKS = (2, 3, 4, 6, 8)
STACKS = (8, 12, 16)
LOOPS = ((1, 1), (1, 0), (0, 0))   # (while_while, refill)
PIPELINES = (1, 2, 3, 4)
TAILS = (0, 3, 4, 5)   # path completion from this bounce (0: off), where below the bounce limit


def cache_path():
    return Path(os.environ.get("RT_AUTOTUNE_CACHE", Path.home() / ".cache" / "rt_amd" / "autotune.json"))


def scene_key(n_tris, width, height, max_bounces):
    return f"tris2^{int(round(math.log2(max(n_tris, 1))))}-px{width}x{height}-b{max_bounces}"


def load_cache():
    p = cache_path()
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def store_cache(cache):
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(cache, indent=1, sort_keys=True))
    except OSError:
        pass


def apply(session, cfg):
    for key in ("blocks_per_cu", "short_stack", "refill", "while_while", "pipelines", "tail_bounce"):
        if key not in cfg:
            continue
        if not session.set_option(key, cfg[key]):
            raise ValueError(f"option {key}={cfg[key]} rejected")


def measure(session, camera, settings, frames, repeats=2):
    """Best of `repeats` timed runs of `frames` frames each, after one warm frame (option
    changes re-size buffers on the next frame): single short runs are too noisy to rank
    configurations a few percent apart."""
    session.render_resident(camera, settings)          # warm (and reset accumulation)
    best = 0.0
    for _ in range(repeats):
        rays, t0 = 0, time.perf_counter()
        for _ in range(frames):
            session.render_resident(camera, settings)
            c, s = session.ray_counts()
            rays += c + s
        best = max(best, rays / (time.perf_counter() - t0) / 1e6)
    return best


def autotune(triangles, camera, settings, width, height, frames=6, widths=WIDTHS, use_cache=True, lib=None, log=None,
             max_leaf=MAX_LEAF):
    """Returns (best config dict, list of all measurements)."""
    key = scene_key(len(triangles), width, height, settings.maxBounces)
    cache = load_cache() if use_cache else {}
    if use_cache and key in cache:
        return cache[key], []
    results = []
    best = None
    configs = [(w, ml) for w in widths for ml in max_leaf]

    def trial(s, cfg):
        nonlocal best
        apply(s, cfg)
        cfg["mrays_per_s"] = round(measure(s, camera, settings, frames), 2)
        results.append(cfg)
        if log:
            log(f"autotune {cfg}")
        if best is None or cfg["mrays_per_s"] > best["mrays_per_s"]:
            best = dict(cfg)

    for w, ml in configs:
        with Session(lib) as s:
            s.set_option("bvh_width", w)
            s.set_option("build_max_leaf", ml)
            s.initialize(triangles, width, height)
            for b in BLOCKS:
                for st in STACKS:
                    for ww, r in LOOPS:
                        trial(s, {"bvh_width": w, "build_max_leaf": ml, "blocks_per_cu": b, "short_stack": st,
                                  "while_while": ww, "refill": r, "pipelines": 2, "tail_bounce": 0})
    with Session(lib) as s:
        s.set_option("bvh_width", best["bvh_width"])
        s.set_option("build_max_leaf", best["build_max_leaf"])
        s.initialize(triangles, width, height)
        launch = {k: v for k, v in best.items() if k != "mrays_per_s"}
        for p in PIPELINES:
            if p != launch["pipelines"]:
                trial(s, dict(launch, pipelines=p))
        launch = {k: v for k, v in best.items() if k != "mrays_per_s"}
        for b in BLOCKS:   # the best workgroup count depends on how many pipelines share the GPU
            if b != launch["blocks_per_cu"]:
                trial(s, dict(launch, blocks_per_cu=b))
        launch = {k: v for k, v in best.items() if k != "mrays_per_s"}
        for t in TAILS:
            if t != launch.get("tail_bounce", 0) and t < settings.maxBounces:
                trial(s, dict(launch, tail_bounce=t))
        # The stack and loop ranking shifts with the pipelines and workgroups that share the
        # GPU: sweep them again at the chosen ones.
        launch = {k: v for k, v in best.items() if k != "mrays_per_s"}
        for st in STACKS:
            for ww, r in LOOPS:
                if (st, ww, r) != (launch["short_stack"], launch["while_while"], launch["refill"]):
                    trial(s, dict(launch, short_stack=st, while_while=ww, refill=r))
    if use_cache:
        cache[key] = best
        store_cache(cache)
    return best, results


def measure_batch(views, cfg, frames=4, repeats=2, lib=None):
    """Rays per second of a batch built with launch options cfg: its frames are captured as
    one graph per frame at build time, so each configuration is a fresh build."""
    with Batch(lib, **cfg) as b:
        for t, w, h, cam, st in views:
            b.add_view(t, w, h, cam, st)
        b.build()
        b.render(1)   # warm
        best = 0.0
        for _ in range(repeats):
            rays, t0 = 0, time.perf_counter()
            for _ in range(frames):
                b.render(1)
                c, s = b.ray_counts()
                rays += c + s
            best = max(best, rays / (time.perf_counter() - t0) / 1e6)
    return best


def autotune_batch(views, widths=WIDTHS, lib=None, log=None):
    """Launch options of a batch of views (BVH width, LDS stack, loop, path pipelines,
    persistent workgroups per CU), searched as autotune() does: the width x stack x loop grid at
    2 pipelines and 3 workgroups per CU, then pipelines, workgroups, and stack x loop again.
    Returns (best config dict, list of all measurements)."""
    results, best = [], None

    def trial(cfg):
        nonlocal best
        cfg = dict(cfg)
        cfg["mrays_per_s"] = round(measure_batch(views, cfg, lib=lib), 2)
        results.append(cfg)
        if log:
            log(f"autotune batch {cfg}")
        if best is None or cfg["mrays_per_s"] > best["mrays_per_s"]:
            best = dict(cfg)

    for w in widths:
        for st in STACKS:
            for ww, r in LOOPS:
                trial({"bvh_width": w, "blocks_per_cu": 3, "short_stack": st, "while_while": ww, "refill": r,
                       "pipelines": 2})
    launch = {k: v for k, v in best.items() if k != "mrays_per_s"}
    for p in PIPELINES:
        if p != launch["pipelines"]:
            trial(dict(launch, pipelines=p))
    launch = {k: v for k, v in best.items() if k != "mrays_per_s"}
    for b in BLOCKS:
        if b != launch["blocks_per_cu"]:
            trial(dict(launch, blocks_per_cu=b))
    launch = {k: v for k, v in best.items() if k != "mrays_per_s"}
    for st in STACKS:
        for ww, r in LOOPS:
            if (st, ww, r) != (launch["short_stack"], launch["while_while"], launch["refill"]):
                trial(dict(launch, short_stack=st, while_while=ww, refill=r))
    return best, results

