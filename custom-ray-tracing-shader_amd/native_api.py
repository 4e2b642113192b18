"""ctypes binding of libCudaRayTracer.so, the MI355X path tracer's plain C ABI.

Mirrors the reference's Tools/native_api.py:1-48 (same struct names and layouts, same
``load_library`` / ``unpack`` helpers and the ``RAYTRACER_LIBRARY`` override), so scripts
written against the reference's ABI drive this library unchanged.  The extension exports
declared in include/rt_abi.h get their argument types here too.
"""
import ctypes as c
import os
from pathlib import Path

import numpy as np


class Vector(c.Structure):
    _fields_ = [(axis, c.c_float) for axis in "xyz"]


class Triangle(c.Structure):
    _fields_ = [(name, Vector) for name in ("origin", "edge1", "edge2", "color")]
    _fields_ += [("emission", c.c_float), ("oneSided", c.c_int)]
    _fields_ += [(name, Vector) for name in ("normal0", "normal1", "normal2")]
    _fields_ += [("mesh", c.c_int), ("specularColor", Vector), ("smoothness", c.c_float),
                 ("specularProbability", c.c_float)]


class Camera(c.Structure):
    _fields_ = [(name, Vector) for name in ("origin", "right", "up", "forward")]
    _fields_ += [("tanHalfFov", c.c_float), ("aspect", c.c_float), ("orthographicSize", c.c_float),
                 ("orthographic", c.c_int)]


class Settings(c.Structure):
    _fields_ = [("maxBounces", c.c_int)]
    _fields_ += [(name, c.c_float) for name in ("exposure", "aperture", "focusDistance", "skyStrength")]
    _fields_ += [("seed", c.c_uint), ("raysPerPixel", c.c_int)]


class Ray(c.Structure):
    _fields_ = [("origin", c.c_float * 3), ("limit", c.c_float), ("direction", c.c_float * 3), ("flags", c.c_int)]


class Hit(c.Structure):
    _fields_ = [("index", c.c_int), ("distance", c.c_float), ("u", c.c_float), ("v", c.c_float)]


class BvhNode(c.Structure):
    _fields_ = [("low", c.c_float * 3), ("first", c.c_int), ("high", c.c_float * 3), ("count", c.c_int)]


assert (c.sizeof(Triangle), c.sizeof(Camera), c.sizeof(Settings)) == (116, 64, 28)
assert (c.sizeof(Ray), c.sizeof(Hit), c.sizeof(BvhNode)) == (32, 16, 32)

# numpy views of the same layouts, for building large scenes without per-triangle objects.
VEC = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4")])
TRIANGLE_DTYPE = np.dtype([("origin", VEC), ("edge1", VEC), ("edge2", VEC), ("color", VEC),
                           ("emission", "<f4"), ("oneSided", "<i4"), ("normal0", VEC), ("normal1", VEC),
                           ("normal2", VEC), ("mesh", "<i4"), ("specularColor", VEC),
                           ("smoothness", "<f4"), ("specularProbability", "<f4")])
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("limit", "<f4"), ("direction", "<f4", 3), ("flags", "<i4")])
HIT_DTYPE = np.dtype([("index", "<i4"), ("distance", "<f4"), ("u", "<f4"), ("v", "<f4")])
BVH_NODE_DTYPE = np.dtype([("low", "<f4", 3), ("first", "<i4"), ("high", "<f4", 3), ("count", "<i4")])
assert TRIANGLE_DTYPE.itemsize == 116 and RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16

PACKAGE = Path(__file__).resolve().parent
LIBRARY = PACKAGE / "libCudaRayTracer.so"
PIXEL_SUM_FLOATS = 10  # radiance xyz, albedo xyz, normal xyz, covered


def load_library(path=None):
    """Load the plugin (``RAYTRACER_LIBRARY`` overrides the in-tree build)."""
    path = path or os.environ.get("RAYTRACER_LIBRARY", LIBRARY)
    if not Path(path).exists():
        raise FileNotFoundError(f"{path} is not built; run __graft_entry__.build() or make -C {PACKAGE}")
    lib = c.CDLL(str(Path(path).resolve()))
    vp = c.c_void_p
    lib.RtCreate.restype = vp
    lib.RtCreate.argtypes = []
    for name in ("RtDestroy", "RtShutdown", "RtReset"):
        getattr(lib, name).argtypes = [vp]
        getattr(lib, name).restype = None
    lib.RtSampleCount.argtypes = [vp]
    lib.RtSampleCount.restype = c.c_uint
    lib.RtAccumulationVersion.argtypes = [vp]
    lib.RtAccumulationVersion.restype = c.c_uint
    lib.RtCpuThreadCount.argtypes = [vp]
    lib.RtEmbreeBuildMilliseconds.argtypes = [vp]
    lib.RtEmbreeBuildMilliseconds.restype = c.c_float
    lib.RtBvhInfo.argtypes = [vp, c.POINTER(c.c_int), c.POINTER(c.c_int), c.POINTER(c.c_float)]
    lib.RtBvhInfo.restype = None
    lib.RtLastError.argtypes = [vp]
    lib.RtLastError.restype = c.c_char_p
    lib.RtInitialize.argtypes = [vp, c.POINTER(Triangle), c.c_int, c.c_int, c.c_int]
    lib.RtRender.argtypes = [vp, c.POINTER(Camera), c.POINTER(Settings), c.POINTER(c.c_ubyte), c.c_int, c.c_int,
                             c.POINTER(c.c_float)]
    lib.RtDenoise.argtypes = [vp, c.POINTER(c.c_ubyte), c.c_int, c.POINTER(c.c_float)]
    lib.RtTraceRays.argtypes = [vp, c.POINTER(Ray), c.c_int, c.c_int, c.POINTER(Hit)]
    lib.RtTraceRaysNoAlloc.argtypes = [vp, vp, c.c_int, c.c_int, vp]   # device rays, device hits
    lib.RtSobolBatch.argtypes = [c.POINTER(c.c_uint), c.POINTER(c.c_uint), c.c_int, c.c_uint, c.c_int,
                                 c.POINTER(c.c_float)]
    lib.RtBvhExport.argtypes = [vp, c.POINTER(BvhNode), c.c_int, c.POINTER(c.c_int), c.c_int]
    lib.RtReadAccumulation.argtypes = [vp, c.POINTER(c.c_float), c.c_longlong]
    lib.RtRayCounts.argtypes = [vp, c.POINTER(c.c_ulonglong), c.POINTER(c.c_ulonglong)]
    lib.RtSetOption.argtypes = [vp, c.c_char_p, c.c_int]
    lib.RtGetOption.argtypes = [vp, c.c_char_p]
    lib.RtRenderResident.argtypes = [vp, c.POINTER(Camera), c.POINTER(Settings), c.POINTER(c.c_float)]
    lib.RtAutotune.argtypes = [vp, c.POINTER(Camera), c.POINTER(Settings), c.c_int, c.POINTER(c.c_float),
                               c.POINTER(c.c_int)]
    lib.RtCopyOutput.argtypes = [vp, c.POINTER(c.c_ubyte), c.c_int]
    lib.RtStageTimes.argtypes = [vp, c.POINTER(c.c_float), c.POINTER(c.c_int), c.c_int]
    lib.RtStageBusy.argtypes = [vp, c.c_uint, c.POINTER(c.c_float)]
    lib.RtTraceStats.argtypes = [vp, c.POINTER(c.c_ulonglong)]
    lib.RtBandwidthProbe.argtypes = [c.c_longlong, c.c_int, c.POINTER(c.c_float)]
    fp = c.POINTER(c.c_float)
    lib.RtSetDenoiser.argtypes = [vp, vp]
    lib.RtDenoiseStats.argtypes = [vp, c.POINTER(c.c_int), c.POINTER(c.c_int), fp]
    lib.RtReadDenoiserInputs.argtypes = [vp, fp, c.c_longlong]
    lib.RtDenoiserCreate.restype = vp
    lib.RtDenoiserCreate.argtypes = []
    lib.RtDenoiserDestroy.argtypes = [vp]
    lib.RtDenoiserDestroy.restype = None
    lib.RtDenoiserParamCount.argtypes = [vp]
    lib.RtDenoiserParamCount.restype = c.c_longlong
    lib.RtDenoiserInit.argtypes = [vp, c.c_uint]
    lib.RtDenoiserGetParams.argtypes = [vp, fp]
    lib.RtDenoiserSetParams.argtypes = [vp, fp]
    lib.RtDenoiserSave.argtypes = [vp, c.c_char_p]
    lib.RtDenoiserLoad.argtypes = [vp, c.c_char_p]
    lib.RtDenoiserInfer.argtypes = [vp, fp, c.c_int, c.c_int, c.c_int, fp]
    lib.RtDenoiserGradients.argtypes = [vp, fp, fp, c.c_int, c.c_int, c.c_int, fp, fp]
    lib.RtDenoiserSetDataset.argtypes = [vp, fp, fp, c.c_int, c.c_int, c.c_int]
    lib.RtDenoiserTrain.argtypes = [vp, c.c_int, c.c_int, c.c_int, c.c_float, c.c_uint, fp]
    lib.RtDenoiserTimings.argtypes = [vp, fp, fp]
    lib.RtDenoiserSetDisplayLoss.argtypes = [vp, c.c_float]
    lib.RtDenoiserForwardFlops.argtypes = [c.c_int, c.c_int, c.c_int]
    lib.RtDenoiserForwardFlops.restype = c.c_double
    lib.RtConv3x3Forward.argtypes = [fp, c.c_int, c.c_int, c.c_int, c.c_int, fp, fp, c.c_int, c.c_int, fp]
    lib.RtConv3x3Backward.argtypes = [fp, c.c_int, c.c_int, c.c_int, c.c_int, fp, fp, c.c_int, fp, fp, fp]
    lib.RtBatchCreate.restype = vp
    lib.RtBatchCreate.argtypes = []
    lib.RtBatchDestroy.argtypes = [vp]
    lib.RtBatchDestroy.restype = None
    lib.RtBatchLastError.argtypes = [vp]
    lib.RtBatchLastError.restype = c.c_char_p
    lib.RtBatchAddView.argtypes = [vp, c.POINTER(Triangle), c.c_int, c.c_int, c.c_int, c.POINTER(Camera),
                                   c.POINTER(Settings)]
    lib.RtBatchSetOption.argtypes = [vp, c.c_char_p, c.c_int]
    lib.RtBatchBuild.argtypes = [vp, fp]
    lib.RtBatchGetOption.argtypes = [vp, c.c_char_p]
    lib.RtBatchAutotune.argtypes = [vp, c.c_int, fp, c.POINTER(c.c_int)]
    lib.RtBatchRender.argtypes = [vp, c.c_int, fp]
    lib.RtBatchReset.argtypes = [vp]
    lib.RtBatchRayCounts.argtypes = [vp, c.POINTER(c.c_ulonglong), c.POINTER(c.c_ulonglong)]
    lib.RtBatchViewInfo.argtypes = [vp, c.c_int, c.POINTER(c.c_int), c.POINTER(c.c_int), c.POINTER(c.c_uint)]
    lib.RtBatchReadView.argtypes = [vp, c.c_int, c.POINTER(c.c_ubyte), c.c_int, fp, c.c_longlong]
    return lib


class Batch:
    """Batched views rendered together (RtBatch* exports)."""

    def __init__(self, lib=None, **options):
        self.lib = lib or load_library()
        self.handle = self.lib.RtBatchCreate()
        self.sizes = []
        for k, v in options.items():
            if not self.lib.RtBatchSetOption(self.handle, k.encode(), int(v)):
                raise ValueError(f"batch option {k}={v} rejected")

    def close(self):
        if self.handle:
            self.lib.RtBatchDestroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def check(self, ok):
        if not ok:
            raise RuntimeError(self.lib.RtBatchLastError(self.handle).decode())

    def add_view(self, triangles, width, height, camera, settings):
        self.check(self.lib.RtBatchAddView(self.handle, triangle_pointer(triangles), len(triangles), width, height,
                                           c.byref(camera), c.byref(settings)))
        self.sizes.append((width, height))

    def build(self):
        ms = c.c_float()
        self.check(self.lib.RtBatchBuild(self.handle, c.byref(ms)))
        return ms.value

    def render(self, frames=1):
        ms = c.c_float()
        self.check(self.lib.RtBatchRender(self.handle, frames, c.byref(ms)))
        return ms.value

    def reset(self):
        self.check(self.lib.RtBatchReset(self.handle))

    def ray_counts(self):
        a, b = c.c_ulonglong(), c.c_ulonglong()
        self.check(self.lib.RtBatchRayCounts(self.handle, c.byref(a), c.byref(b)))
        return a.value, b.value

    def read(self, view):
        w, h = self.sizes[view]
        rgba = (c.c_ubyte * (w * h * 4))()
        acc = np.zeros(w * h * 10, np.float32)
        self.check(self.lib.RtBatchReadView(self.handle, view, rgba, len(rgba), as_pointer(acc, c.c_float), acc.size))
        return (np.frombuffer(bytes(rgba), dtype=np.uint8).reshape(h, w, 4), acc.reshape(h, w, 10))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Denoiser:
    """The learned denoiser network (RtDenoiser* exports)."""

    def __init__(self, lib=None, seed=None, weights=None):
        self.lib = lib or load_library()
        self.handle = self.lib.RtDenoiserCreate()
        if not self.handle:
            raise RuntimeError("RtDenoiserCreate failed")
        self.count = self.lib.RtDenoiserParamCount(self.handle)
        if weights is not None:
            if not self.lib.RtDenoiserLoad(self.handle, str(weights).encode()):
                raise RuntimeError(f"cannot load {weights}")
        elif seed is not None:
            assert self.lib.RtDenoiserInit(self.handle, seed)

    def close(self):
        if self.handle:
            self.lib.RtDenoiserDestroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def params(self):
        out = np.zeros(self.count, np.float32)
        assert self.lib.RtDenoiserGetParams(self.handle, as_pointer(out, c.c_float))
        return out

    @params.setter
    def params(self, value):
        value = _f32(value)
        assert value.size == self.count and self.lib.RtDenoiserSetParams(self.handle, as_pointer(value, c.c_float))

    def save(self, path):
        assert self.lib.RtDenoiserSave(self.handle, str(path).encode())

    def infer(self, raw):
        raw = _f32(raw)
        n, h, w, _ = raw.shape
        out = np.zeros((n, h, w, 3), np.float32)
        assert self.lib.RtDenoiserInfer(self.handle, as_pointer(raw, c.c_float), n, h, w, as_pointer(out, c.c_float))
        return out

    def gradients(self, raw, target):
        raw, target = _f32(raw), _f32(target)
        n, h, w, _ = raw.shape
        loss = c.c_float()
        grads = np.zeros(self.count, np.float32)
        assert self.lib.RtDenoiserGradients(self.handle, as_pointer(raw, c.c_float), as_pointer(target, c.c_float), n, h,
                                            w, c.byref(loss), as_pointer(grads, c.c_float))
        return loss.value, grads

    def set_dataset(self, inputs, targets):
        inputs, targets = _f32(inputs), _f32(targets)
        n, h, w, _ = inputs.shape
        assert self.lib.RtDenoiserSetDataset(self.handle, as_pointer(inputs, c.c_float), as_pointer(targets, c.c_float),
                                             n, h, w)

    def train(self, steps, batch=8, crop=128, lr=2e-4, seed=1):
        losses = np.zeros(steps, np.float32)
        assert self.lib.RtDenoiserTrain(self.handle, steps, batch, crop, lr, seed, as_pointer(losses, c.c_float))
        return losses

    def timings(self):
        f, t = c.c_float(), c.c_float()
        self.lib.RtDenoiserTimings(self.handle, c.byref(f), c.byref(t))
        return f.value, t.value

    def set_display_loss(self, weight):
        if not self.lib.RtDenoiserSetDisplayLoss(self.handle, float(weight)):
            raise ValueError(f"display loss weight {weight} rejected")


STAGES = ("reset", "raygen", "closest", "shade", "shadow", "accumulate", "tail")


def unpack(structure, values):
    """Read Unity's JSON snapshot without changing its float inputs (Tools/native_api.py:45-48)."""
    return structure(**{name: Vector(**values[name]) if field is Vector else values[name]
                        for name, field in structure._fields_})


def triangle_pointer(array):
    """ctypes pointer to a contiguous TRIANGLE_DTYPE array (or a ctypes Triangle array)."""
    if isinstance(array, np.ndarray):
        assert array.dtype == TRIANGLE_DTYPE and array.flags["C_CONTIGUOUS"]
        return array.ctypes.data_as(c.POINTER(Triangle))
    return c.cast(array, c.POINTER(Triangle))


def as_pointer(array, ctype):
    return array.ctypes.data_as(c.POINTER(ctype))


class Session:
    """Small RAII wrapper over one plugin session (RtCreate ... RtDestroy)."""

    def __init__(self, lib=None):
        self.lib = lib or load_library()
        self.handle = self.lib.RtCreate()
        self.width = self.height = 0

    def close(self):
        if self.handle:
            self.lib.RtDestroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def error(self):
        return self.lib.RtLastError(self.handle).decode()

    def check(self, ok):
        if not ok:
            raise RuntimeError(self.error())

    def set_option(self, key, value):
        return self.lib.RtSetOption(self.handle, key.encode(), int(value))

    def get_option(self, key):
        return self.lib.RtGetOption(self.handle, key.encode())

    def initialize(self, triangles, width, height):
        n = len(triangles)
        self.check(self.lib.RtInitialize(self.handle, triangle_pointer(triangles), n, width, height))
        self.width, self.height = width, height

    def render(self, camera, settings, backend=2):
        out = (c.c_ubyte * (self.width * self.height * 4))()
        ms = c.c_float()
        self.check(self.lib.RtRender(self.handle, c.byref(camera), c.byref(settings), out, len(out), backend,
                                     c.byref(ms)))
        return np.frombuffer(bytes(out), dtype=np.uint8).reshape(self.height, self.width, 4), ms.value

    def render_resident(self, camera, settings):
        """GPU frame left in device memory; returns the kernels' device milliseconds."""
        ms = c.c_float()
        self.check(self.lib.RtRenderResident(self.handle, c.byref(camera), c.byref(settings), c.byref(ms)))
        return ms.value

    def copy_output(self):
        out = (c.c_ubyte * (self.width * self.height * 4))()
        self.check(self.lib.RtCopyOutput(self.handle, out, len(out)))
        return np.frombuffer(bytes(out), dtype=np.uint8).reshape(self.height, self.width, 4)

    def stage_times(self):
        ms = (c.c_float * len(STAGES))()
        n = (c.c_int * len(STAGES))()
        self.check(self.lib.RtStageTimes(self.handle, ms, n, len(STAGES)))
        return {k: (ms[i], n[i]) for i, k in enumerate(STAGES)}

    def stage_busy(self, *stages):
        """Milliseconds of the last frame during which a launch of any of the named stages ran."""
        ms = c.c_float()
        self.check(self.lib.RtStageBusy(self.handle, sum(1 << STAGES.index(k) for k in stages), c.byref(ms)))
        return ms.value

    def trace_stats(self):
        v = (c.c_ulonglong * 4)()
        self.check(self.lib.RtTraceStats(self.handle, v))
        return {"closest_nodes": v[0], "closest_tris": v[1], "shadow_nodes": v[2], "shadow_tris": v[3]}

    def accumulation(self):
        out = np.zeros(self.width * self.height * PIXEL_SUM_FLOATS, dtype=np.float32)
        self.check(self.lib.RtReadAccumulation(self.handle, as_pointer(out, c.c_float), out.size))
        return out.reshape(self.height, self.width, PIXEL_SUM_FLOATS)

    def ray_counts(self):
        a, b = c.c_ulonglong(), c.c_ulonglong()
        self.check(self.lib.RtRayCounts(self.handle, c.byref(a), c.byref(b)))
        return a.value, b.value

    def trace(self, rays, any_hit=False):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(len(rays), dtype=HIT_DTYPE)
        self.check(self.lib.RtTraceRays(self.handle, as_pointer(rays, Ray), len(rays), int(any_hit),
                                        as_pointer(hits, Hit)))
        return hits

    def bvh(self):
        count = self.lib.RtBvhExport(self.handle, None, 0, None, 0)
        nodes = np.zeros(count, dtype=BVH_NODE_DTYPE)
        nodes_n = self.lib.RtBvhExport(self.handle, as_pointer(nodes, BvhNode), count, None, 0)
        assert nodes_n == count
        return nodes

    def bvh_order(self, n):
        order = np.zeros(n, dtype=np.int32)
        self.lib.RtBvhExport(self.handle, None, 0, as_pointer(order, c.c_int), n)
        return order

    def denoise(self):
        out = (c.c_ubyte * (self.width * self.height * 4))()
        ms = c.c_float()
        self.check(self.lib.RtDenoise(self.handle, out, len(out), c.byref(ms)))
        return np.frombuffer(bytes(out), dtype=np.uint8).reshape(self.height, self.width, 4), ms.value

    def denoise_stats(self):
        t, r, ms = c.c_int(), c.c_int(), c.c_float()
        self.lib.RtDenoiseStats(self.handle, c.byref(t), c.byref(r), c.byref(ms))
        return t.value, r.value, ms.value

    def denoiser_inputs(self):
        out = np.zeros(self.width * self.height * 10, np.float32)
        self.check(self.lib.RtReadDenoiserInputs(self.handle, as_pointer(out, c.c_float), out.size))
        return out.reshape(self.height, self.width, 10)

    def bvh_info(self):
        nodes, depth, ms = c.c_int(), c.c_int(), c.c_float()
        self.lib.RtBvhInfo(self.handle, c.byref(nodes), c.byref(depth), c.byref(ms))
        return nodes.value, depth.value, ms.value


def sobol_batch(lib, pixels, samples, seed, dims):
    pixels = np.ascontiguousarray(pixels, dtype=np.uint32)
    samples = np.ascontiguousarray(samples, dtype=np.uint32)
    out = np.zeros((len(pixels), dims), dtype=np.float32)
    ok = lib.RtSobolBatch(as_pointer(pixels, c.c_uint), as_pointer(samples, c.c_uint), len(pixels), seed, dims,
                          as_pointer(out, c.c_float))
    if not ok:
        raise RuntimeError("RtSobolBatch failed")
    return out
