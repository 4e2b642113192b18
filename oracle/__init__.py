"""CPU oracle of the path-tracing hot path - TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import this
package, and only as the checker.  The product library never links or calls it.

The C++ part (oracle.cpp, built by the Makefile here into build/liboracle.so) restates the
reference's algorithms; this module loads it and adds the Sobol direction table, computed
here from Joe & Kuo's published new-joe-kuo-6.21201 data (SciPy ships that file).

Function -> reference -> what pins it (besides itself):
  sobol_table        Native/sampling.cuh:7-15    rocRAND JOEKUO6 table equality
  hash / sobol_batch Native/sampling.cuh:21-35   hand-computed hashes; stratification
                                                 (Tools/verify_sampling.cu:16-25); golden PNGs
  intersect32/64     Native/renderer.cu:99-112   closed-form hits, parallel/one-sided/degenerate
                                                 cases, FP32 vs FP64 agreement
  intersect_wt32     Woop et al. 2013 (watertight) closed-form hits, FP64 agreement, no gaps at
                                                 shared edges/vertices of triangle fans
  build_bvh / trace  Native/bvh.cuh:35-105,      structural invariants + brute-force
                     Native/renderer.cu:136-178  differential (Tools/verify_bvh.cu:9-66)
  sah_cost           Native/bvh.cuh:66,89        hand-built trees with known areas
  OracleScene.render Native/renderer.cu:197-302  verify_native known answers and the
                                                 reference's golden CUDA PNGs
"""
import ctypes as c
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# RT_ORACLE_LIB selects another build (the sanitizer build, make SAN=1: build_san/liboracle.so).
LIB_PATH = Path(os.environ.get("RT_ORACLE_LIB", HERE / "build" / "liboracle.so"))

VEC = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4")])
TRIANGLE = np.dtype([("origin", VEC), ("edge1", VEC), ("edge2", VEC), ("color", VEC), ("emission", "<f4"),
                     ("oneSided", "<i4"), ("normal0", VEC), ("normal1", VEC), ("normal2", VEC), ("mesh", "<i4"),
                     ("specularColor", VEC), ("smoothness", "<f4"), ("specularProbability", "<f4")])
RAY = np.dtype([("origin", "<f4", 3), ("limit", "<f4"), ("direction", "<f4", 3), ("flags", "<i4")])
HIT = np.dtype([("index", "<i4"), ("distance", "<f4"), ("u", "<f4"), ("v", "<f4")])
NODE = np.dtype([("low", "<f4", 3), ("first", "<i4"), ("high", "<f4", 3), ("count", "<i4")])
CAMERA = np.dtype([("origin", VEC), ("right", VEC), ("up", VEC), ("forward", VEC), ("tanHalfFov", "<f4"),
                   ("aspect", "<f4"), ("orthographicSize", "<f4"), ("orthographic", "<i4")])
SETTINGS = np.dtype([("maxBounces", "<i4"), ("exposure", "<f4"), ("aperture", "<f4"), ("focusDistance", "<f4"),
                     ("skyStrength", "<f4"), ("seed", "<u4"), ("raysPerPixel", "<i4")])
assert (TRIANGLE.itemsize, CAMERA.itemsize, SETTINGS.itemsize) == (116, 64, 28)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = c.CDLL(str(LIB_PATH))
        P = c.c_void_p
        L.or_hash.argtypes = [c.c_uint]
        L.or_hash.restype = c.c_uint
        L.or_sobol_batch.argtypes = [P, P, P, c.c_int, c.c_uint, c.c_int, P]
        L.or_intersect32.argtypes = [P, P, P, P, P]
        L.or_intersect32.restype = c.c_float
        L.or_intersect64.argtypes = [P, P, P, P, P]
        L.or_intersect64.restype = c.c_double
        L.or_brute32.argtypes = [P, c.c_int, P, c.c_int, c.c_int, P]
        L.or_brute64.argtypes = [P, c.c_int, P, c.c_int, P, P]
        L.or_brute_wt32.argtypes = [P, c.c_int, P, c.c_int, c.c_int, P]
        L.or_intersect_wt32.argtypes = [P, P, P, P, P]
        L.or_intersect_wt32.restype = c.c_float
        L.or_bvh_build.argtypes = [P, c.c_int, c.c_int, c.c_int]
        L.or_bvh_build.restype = P
        L.or_bvh_import.argtypes = [P, c.c_int, P, c.c_int]
        L.or_bvh_import.restype = P
        L.or_bvh_free.argtypes = [P]
        L.or_bvh_node_count.argtypes = [P]
        L.or_bvh_depth.argtypes = [P]
        L.or_bvh_export.argtypes = [P, P, P]
        L.or_trace_rays.argtypes = [P, P, P, c.c_int, c.c_int, P, P, P]
        L.or_sah_cost.argtypes = [P, c.c_int]
        L.or_sah_cost.restype = c.c_double
        L.or_scene_create.argtypes = [P, c.c_int, P]
        L.or_scene_create.restype = P
        L.or_scene_free.argtypes = [P]
        L.or_scene_trace_rays.argtypes = [P, P, c.c_int, c.c_int, P]
        L.or_scene_bvh_nodes.argtypes = [P]
        L.or_scene_sah_cost.argtypes = [P]
        L.or_scene_sah_cost.restype = c.c_double
        L.or_scene_bvh_depth.argtypes = [P]
        L.or_scene_light_area.argtypes = [P]
        L.or_scene_light_area.restype = c.c_float
        L.or_render.argtypes = [P, P, P, c.c_int, c.c_int, c.c_uint, P, P, c.c_int, c.c_int, c.c_int, P]
        L.or_display.argtypes = [c.c_float, c.c_float]
        L.or_display.restype = c.c_ubyte
        L.or_srgb_decode.argtypes = [c.c_float]
        L.or_srgb_decode.restype = c.c_float
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(c.c_void_p)


# --------------------------------------------------------------------------- Sobol table
def sobol_table(ndim=260):
    """Joe-Kuo 6.21201 direction numbers (dimension 0 = van der Corput), shape (ndim, 32)."""
    import scipy
    data = np.load(os.path.join(os.path.dirname(scipy.__file__), "stats", "_sobol_direction_numbers.npz"))
    poly, vinit = data["poly"], data["vinit"]
    out = np.zeros((ndim, 32), dtype=np.uint32)
    out[0] = [1 << (31 - k) for k in range(32)]
    for j in range(1, ndim):
        a = int(poly[j])
        s = a.bit_length() - 1           # polynomial degree
        m = [int(vinit[j, k]) for k in range(s)]
        v = [0] * 32
        for k in range(32):
            if k < s:
                v[k] = m[k] << (31 - k)
                continue
            val = v[k - s] ^ (v[k - s] >> s)
            for bit in range(1, s):       # inner coefficients a_1 .. a_{s-1}
                if (a >> (s - bit)) & 1:
                    val ^= v[k - bit]
            v[k] = val & 0xFFFFFFFF
        out[j] = v
    return out


_TABLE = None


def directions():
    global _TABLE
    if _TABLE is None:
        _TABLE = np.ascontiguousarray(sobol_table(260).reshape(-1))
    return _TABLE


def hash32(x):
    return int(lib().or_hash(x))


def sobol_batch(pixels, samples, seed, dims):
    pixels = np.ascontiguousarray(pixels, dtype=np.uint32)
    samples = np.ascontiguousarray(samples, dtype=np.uint32)
    out = np.zeros((len(pixels), dims), dtype=np.float32)
    lib().or_sobol_batch(_p(directions()), _p(pixels), _p(samples), len(pixels), seed, dims, _p(out))
    return out


# --------------------------------------------------------------------------- geometry
def intersect32(tri, o, d):
    tri = np.ascontiguousarray(tri, dtype=TRIANGLE).reshape(1)
    o = np.asarray(o, dtype=np.float32); d = np.asarray(d, dtype=np.float32)
    u, v = np.zeros(1, np.float32), np.zeros(1, np.float32)
    t = lib().or_intersect32(_p(tri), _p(o), _p(d), _p(u), _p(v))
    return float(t), float(u[0]), float(v[0])


def intersect64(tri, o, d):
    tri = np.ascontiguousarray(tri, dtype=TRIANGLE).reshape(1)
    o = np.asarray(o, dtype=np.float64); d = np.asarray(d, dtype=np.float64)
    u, v = np.zeros(1), np.zeros(1)
    t = lib().or_intersect64(_p(tri), _p(o), _p(d), _p(u), _p(v))
    return float(t), float(u[0]), float(v[0])


def brute32(tris, rays, any_hit=False):
    tris = np.ascontiguousarray(tris, dtype=TRIANGLE); rays = np.ascontiguousarray(rays, dtype=RAY)
    hits = np.zeros(len(rays), dtype=HIT)
    lib().or_brute32(_p(tris), len(tris), _p(rays), len(rays), int(any_hit), _p(hits))
    return hits


def intersect_wt32(tri, o, d):
    """Watertight test (Woop, Benthin and Wald 2013) with the reference's acceptance rules."""
    tri = np.ascontiguousarray(tri, dtype=TRIANGLE).reshape(1)
    o = np.asarray(o, dtype=np.float32); d = np.asarray(d, dtype=np.float32)
    u, v = np.zeros(1, np.float32), np.zeros(1, np.float32)
    t = lib().or_intersect_wt32(_p(tri), _p(o), _p(d), _p(u), _p(v))
    return float(t), float(u[0]), float(v[0])


def brute_wt32(tris, rays, any_hit=False):
    tris = np.ascontiguousarray(tris, dtype=TRIANGLE); rays = np.ascontiguousarray(rays, dtype=RAY)
    hits = np.zeros(len(rays), dtype=HIT)
    lib().or_brute_wt32(_p(tris), len(tris), _p(rays), len(rays), int(any_hit), _p(hits))
    return hits


def brute64(tris, rays):
    tris = np.ascontiguousarray(tris, dtype=TRIANGLE); rays = np.ascontiguousarray(rays, dtype=RAY)
    idx = np.zeros(len(rays), dtype=np.int32); dist = np.zeros(len(rays), dtype=np.float64)
    lib().or_brute64(_p(tris), len(tris), _p(rays), len(rays), _p(idx), _p(dist))
    return idx, dist


class Bvh:
    """Binary binned-SAH BVH in the reference layout (Native/bvh.cuh)."""

    def __init__(self, tris, bins=12, max_leaf=4, nodes=None, order=None):
        self.tris = np.ascontiguousarray(tris, dtype=TRIANGLE)
        if nodes is None:
            self.h = lib().or_bvh_build(_p(self.tris), len(self.tris), bins, max_leaf)
        else:
            nodes = np.ascontiguousarray(nodes, dtype=NODE); order = np.ascontiguousarray(order, dtype=np.int32)
            self.h = lib().or_bvh_import(_p(nodes), len(nodes), _p(order), len(order))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_bvh_free(self.h)
            self.h = None

    def export(self):
        n = lib().or_bvh_node_count(self.h)
        nodes = np.zeros(n, dtype=NODE); order = np.zeros(len(self.tris), dtype=np.int32)
        lib().or_bvh_export(self.h, _p(nodes), _p(order))
        return nodes, order

    @property
    def depth(self):
        return lib().or_bvh_depth(self.h)

    def trace(self, rays, any_hit=False, counts=False):
        rays = np.ascontiguousarray(rays, dtype=RAY)
        hits = np.zeros(len(rays), dtype=HIT)
        nv, nt = c.c_ulonglong(), c.c_ulonglong()
        lib().or_trace_rays(self.h, _p(self.tris), _p(rays), len(rays), int(any_hit), _p(hits),
                            c.byref(nv), c.byref(nt))
        return (hits, nv.value, nt.value) if counts else hits


def sah_cost(nodes):
    nodes = np.ascontiguousarray(nodes, dtype=NODE)
    return float(lib().or_sah_cost(_p(nodes), len(nodes)))


# --------------------------------------------------------------------------- integrator
class OracleScene:
    """Native/renderer.cu:340-357 preprocessing + the CPU render loop (renderer.cu:289-302,405-416)."""

    def __init__(self, tris):
        self.tris = np.ascontiguousarray(tris, dtype=TRIANGLE)
        self.h = lib().or_scene_create(_p(self.tris), len(self.tris), _p(directions()))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_scene_free(self.h)
            self.h = None

    @property
    def light_area(self):
        return float(lib().or_scene_light_area(self.h))

    @property
    def sah_cost(self):
        """SAH cost of the scene's 12-bin reference build (sah_cost)."""
        return float(lib().or_scene_sah_cost(self.h))

    def trace(self, rays, any_hit=False):
        """Hit records of ray queries against the scene's own BVH (the one render uses)."""
        rays = np.ascontiguousarray(rays, dtype=RAY)
        hits = np.zeros(len(rays), dtype=HIT)
        lib().or_scene_trace_rays(self.h, _p(rays), len(rays), int(any_hit), _p(hits))
        return hits

    def render(self, camera, settings, width, height, sample=0, total=None, threads=0, rows=None):
        """One RtRender call: raysPerPixel samples starting at `sample` added into `total`
        (10 floats per pixel).  Returns (rgba, total, (closest, shadow) ray counts)."""
        camera = np.ascontiguousarray(camera, dtype=CAMERA).reshape(1)
        settings = np.ascontiguousarray(settings, dtype=SETTINGS).reshape(1)
        if total is None:
            total = np.zeros((height, width, 10), dtype=np.float32)
        rgba = np.zeros((height, width, 4), dtype=np.uint8)
        y0, y1 = rows if rows else (0, height)
        rays = np.zeros(2, dtype=np.uint64)
        nt = threads if threads else (os.cpu_count() or 1)
        lib().or_render(self.h, _p(camera), _p(settings), width, height, sample, _p(total), _p(rgba), nt, y0, y1,
                        _p(rays))
        return rgba, total, (int(rays[0]), int(rays[1]))


def display(value, exposure=1.0):
    return int(lib().or_display(value, exposure))


def srgb_decode(value):
    return float(lib().or_srgb_decode(value))


def camera_record(origin, right, up, forward, tan_half_fov, aspect, ortho_size=1.0, orthographic=0):
    cam = np.zeros(1, dtype=CAMERA)
    for name, v in (("origin", origin), ("right", right), ("up", up), ("forward", forward)):
        cam[name]["x"], cam[name]["y"], cam[name]["z"] = v
    cam["tanHalfFov"], cam["aspect"], cam["orthographicSize"], cam["orthographic"] = tan_half_fov, aspect, ortho_size, orthographic
    return cam


def settings_record(max_bounces, exposure=1.0, aperture=0.0, focus=1.0, sky=0.0, seed=42, rpp=1):
    s = np.zeros(1, dtype=SETTINGS)
    s["maxBounces"], s["exposure"], s["aperture"], s["focusDistance"] = max_bounces, exposure, aperture, focus
    s["skyStrength"], s["seed"], s["raysPerPixel"] = sky, seed, rpp
    return s
