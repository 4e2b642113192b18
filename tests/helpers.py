"""Shared test helpers: fixtures, struct conversion and the reference's test scenes."""
import ctypes as c
import math
from pathlib import Path

import numpy as np

import oracle
from native_api import TRIANGLE_DTYPE, Camera, Settings, Vector

GOLDEN = Path(__file__).resolve().parent / "golden"


def cam_record(camera):
    """ctypes Camera -> oracle CAMERA record (same bytes)."""
    return np.frombuffer(bytes(camera), dtype=oracle.CAMERA).copy()


def settings_record(settings):
    return np.frombuffer(bytes(settings), dtype=oracle.SETTINGS).copy()


def suzanne():
    """The reference's Suzanne snapshot and golden CUDA renders (tests/golden/suzanne.npz)."""
    d = np.load(GOLDEN / "suzanne.npz")
    tris = d["triangles"].view(TRIANGLE_DTYPE).copy()
    cv = d["camera"]
    w, h = int(d["width"]), int(d["height"])
    cam = Camera(Vector(*cv[0:3]), Vector(*cv[3:6]), Vector(*cv[6:9]), Vector(*cv[9:12]), float(cv[12]), w / h,
                 float(cv[14]), int(d["orthographic"]))
    st = d["settings"]
    settings = Settings(int(st[0]), float(st[1]), float(st[2]), float(st[3]), float(st[4]), int(st[5]), int(st[6]))
    return tris, cam, settings, w, h, {k: d[k] for k in ("sobol16", "sobol64", "reference2048")}


def image_agreement(a, b):
    """(mean abs channel error, fraction of RGB channels off by more than 1, max error)."""
    d = np.abs(a[..., :3].astype(np.int32) - b[..., :3].astype(np.int32))
    return float(d.mean()), float((d > 1).mean()), int(d.max())


# --- Tools/verify_bvh.cu:35-41,104-139 scenes and rays, restated --------------------------
def vbvh_triangle(a, b, d, sided=0):
    a, b, d = (np.asarray(x, np.float32) for x in (a, b, d))
    t = np.zeros(1, dtype=TRIANGLE_DTYPE)
    e1, e2 = b - a, d - a
    n = np.cross(e1.astype(np.float64), e2.astype(np.float64))
    ln = np.linalg.norm(n)
    n = (n / ln if ln > 0 else n).astype(np.float32)
    for name, v in (("origin", a), ("edge1", e1), ("edge2", e2), ("normal0", n), ("normal1", n), ("normal2", n),
                    ("color", (1, 1, 1)), ("specularColor", (1, 1, 1))):
        t[name]["x"], t[name]["y"], t[name]["z"] = v
    t["oneSided"] = sided
    return t


def mt19937_floats(seed, count, lo=-10.0, hi=10.0):
    """std::uniform_real_distribution<float>(lo, hi) over std::mt19937(seed), as libstdc++ draws it."""
    rs = np.random.RandomState(seed)
    raw = rs._bit_generator.random_raw(count).astype(np.uint32)
    u = (raw.astype(np.float32) / np.float32(4294967296.0)).astype(np.float32)
    u = np.where(u >= 1, np.nextafter(np.float32(1), np.float32(0)), u).astype(np.float32)
    return (u * np.float32(hi - lo) + np.float32(lo)).astype(np.float32)


class _Draws:
    def __init__(self, seed):
        self.buf = mt19937_floats(seed, 400000)
        self.i = 0

    def __call__(self):
        v = self.buf[self.i]
        self.i += 1
        return v


def unit32(v):
    """Native/renderer.cu:82 in float32 scalar arithmetic, left-to-right dot product."""
    x, y, z = (np.float32(t) for t in v)
    dd = np.float32(np.float32(x * x) + np.float32(y * y)) + np.float32(z * z)
    inv = np.float32(1) / np.sqrt(np.maximum(dd, np.float32(1e-20)))
    return np.array([x * inv, y * inv, z * inv], np.float32)


def vbvh_rays_and_scenes():
    """Rays and scenes of Tools/verify_bvh.cu:105-138 (same mt19937(719393) draw order)."""
    draw = _Draws(719393)
    rays = []
    for i in range(12000):
        o = [draw(), draw(), draw()]
        d = unit32([draw(), draw(), draw()])
        if i % 4 == 0:
            d = np.array([(i % 3) == 0, (i % 3) == 1, (i % 3) == 2], np.float32)
        if i % 5 == 0:
            d = d.copy()
            d[0] = np.float32(1e-40)
        limit = np.float32(1e30) if i % 2 else abs(draw())
        rays.append((o, limit, d))
    edge = [vbvh_triangle((-1, -1, 0), (1, -1, 0), (1, 1, 0)), vbvh_triangle((-1, -1, 0), (1, 1, 0), (-1, 1, 0), 1)]
    edge.append(edge[0].copy())
    edge.append(vbvh_triangle((0, 0, 0), (0, 0, 0), (0, 0, 0)))
    for z in (-1.0, 0.0, 1.0):
        for x in (-1.0, 0.0, 1.0):
            for d in ((0, 0, 1), (0, 0, -1), (0, 1, 0), (1e-40, 0, 1)):
                for limit in (.5, 1.0, 2.0, 1e30):
                    rays.append(([x, x, z], np.float32(limit), np.asarray(d, np.float32)))
    scatter = []
    for i in range(3000):
        a = np.array([draw(), draw(), draw()], np.float32)
        scatter.append(vbvh_triangle(a, a + np.float32([.2, .3, 0]), a + np.float32([0, .1, .4]), i % 2))
    deep = []
    for i in range(200):
        x = np.float32(1.5) ** np.float32(i)
        if not np.isfinite(x):
            break
        deep.append(vbvh_triangle((x, -1, 0), (x, 1, 0), (x, 0, 1)))
    ray_arr = np.zeros(len(rays), dtype=oracle.RAY)
    for k, (o, limit, d) in enumerate(rays):
        ray_arr[k]["origin"] = o
        ray_arr[k]["limit"] = limit
        ray_arr[k]["direction"] = d
    scenes = {
        "edges": np.concatenate(edge),
        "scattered": np.concatenate(scatter),
        "overlap": np.concatenate([edge[0]] * 256),
        "unbalanced": np.concatenate(deep),
    }
    return ray_arr, scenes


def check_layout(nodes, order, tris, depth_limit=32):
    """Tools/verify_bvh.cu:42-66: every node once, every triangle in one leaf, nested boxes."""
    n = len(nodes)
    visited = np.zeros(n, np.int32)
    covered = np.zeros(len(tris), np.int32)
    stack = [(0, 0)]
    v0 = np.stack([tris["origin"][a] for a in "xyz"], 1).astype(np.float32)
    v1 = v0 + np.stack([tris["edge1"][a] for a in "xyz"], 1).astype(np.float32)
    v2 = v0 + np.stack([tris["edge2"][a] for a in "xyz"], 1).astype(np.float32)
    while stack:
        i, d = stack.pop()
        assert 0 <= i < n and visited[i] == 0 and d <= depth_limit
        visited[i] += 1
        nd = nodes[i]
        if nd["count"] > 0:
            f, cnt = int(nd["first"]), int(nd["count"])
            assert f >= 0 and f + cnt <= len(order)
            ids = order[f:f + cnt]
            covered[ids] += 1
            for v in (v0[ids], v1[ids], v2[ids]):
                assert (v >= nd["low"]).all() and (v <= nd["high"]).all()
        else:
            for ch in (int(nd["first"]), int(nd["first"]) + 1):
                assert 0 <= ch < n
                assert (nodes[ch]["low"] >= nd["low"]).all() and (nodes[ch]["high"] <= nd["high"]).all()
                stack.append((ch, d + 1))
    assert (visited == 1).all() and (covered == 1).all()


def verify_native_triangle(a, b, d, color=(1, 1, 1), emission=0.0):
    """Tools/verify_native.py:97-103 triangle helper."""
    a, b, d = (np.asarray(x, np.float64) for x in (a, b, d))
    e1, e2 = b - a, d - a
    n = np.cross(e1, e2)
    n = n / math.sqrt((n * n).sum())
    t = np.zeros(1, dtype=TRIANGLE_DTYPE)
    for name, v in (("origin", a), ("edge1", e1), ("edge2", e2), ("color", color), ("normal0", n), ("normal1", n),
                    ("normal2", n), ("specularColor", (1, 1, 1))):
        t[name]["x"], t[name]["y"], t[name]["z"] = v
    t["emission"] = emission
    t["smoothness"] = .7
    return t


def vn_quad(x0, x1, y0, y1, z, **kw):
    return np.concatenate([verify_native_triangle((x0, y0, z), (x1, y0, z), (x1, y1, z), **kw),
                           verify_native_triangle((x0, y0, z), (x1, y1, z), (x0, y1, z), **kw)])
