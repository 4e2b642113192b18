#!/usr/bin/env python3
"""Time the plugin's backends on one Unity scene snapshot, outside Unity.

Restates Tools/benchmark_native.py:19-105: the same command line, the same per-case timing
method (one warm-up call, then ``--passes`` measured calls of RtRender after RtReset; the
median of the library's compute time and of the wall time of the call), the same output
directory (scene copy, raw RGBA per backend, a PNG flipped to top-row-first, results.json)
and the same results.json fields, so compare_benchmarks.py can gate a change on exact image
hashes.  Differences: backend 2 is named "HIP GPU"; a backend that reports itself
unavailable (backend 3 when libembree3.so.3 cannot be loaded) is recorded with its error
instead of aborting the run; no PIL is needed for the PNG.

    python benchmark_native.py scene.json results --sizes 320x180,1280x720 --bounces 3,8
"""
import argparse
import ctypes as c
import hashlib
import json
import os
import statistics
import struct
import sys
import time
import zlib
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
if str(HERE) not in sys.path:
    sys.path.insert(0, str(HERE))

from native_api import LIBRARY, load_library, triangle_pointer  # noqa: E402
from snapshot import load_snapshot  # noqa: E402

BACKEND_NAMES = ("CPU single", "CPU multi", "HIP GPU", "Embree CPU multi")


def write_png(path, rgba, width, height):
    """8-bit RGBA PNG of a bottom-row-first buffer, written top row first."""
    rows = np.frombuffer(rgba, np.uint8).reshape(height, width, 4)[::-1]
    raw = b"".join(b"\x00" + rows[y].tobytes() for y in range(height))

    def chunk(kind, data):
        return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)

    header = struct.pack(">IIBBBBB", width, height, 8, 6, 0, 0, 0)
    Path(path).write_bytes(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", header) + chunk(b"IDAT", zlib.compress(raw, 6)) +
                           chunk(b"IEND", b""))


def channel_errors(reference, image):
    """|a - b| over the RGB channels of two RGBA8 buffers (alpha skipped)."""
    a = np.frombuffer(reference, np.uint8).reshape(-1, 4)[:, :3].astype(np.int32)
    b = np.frombuffer(image, np.uint8).reshape(-1, 4)[:, :3].astype(np.int32)
    return np.abs(a - b).ravel()


def run(snapshot_path, output, library=LIBRARY, backend=None, sizes="320x180,1280x720", bounces="3,8", passes=5,
        log=print):
    output = Path(output)
    if (output / "results.json").exists():
        raise FileExistsError(f"{output / 'results.json'} already exists; give an empty output directory")
    output.mkdir(parents=True, exist_ok=True)
    scene_bytes = Path(snapshot_path).read_bytes()
    (output / "scene.json").write_bytes(scene_bytes)
    snap = load_snapshot(snapshot_path)
    triangles = np.ascontiguousarray(snap.triangles)
    camera, settings = snap.camera, snap.settings
    library = Path(library)
    lib = load_library(library)
    session = lib.RtCreate()
    results = []

    def check(ok):
        if not ok:
            raise RuntimeError(lib.RtLastError(session).decode())

    def write_results():
        doc = {"snapshot_sha256": hashlib.sha256(scene_bytes).hexdigest(), "triangles": len(triangles),
               "library": str(library.resolve()), "library_sha256": hashlib.sha256(library.read_bytes()).hexdigest(),
               "omp_threads": os.getenv("OMP_NUM_THREADS"), "measured_passes": passes, "warmup_passes": 1,
               "method": "Standalone native render calls; same seeded paths per backend, one warm-up followed by "
                         "measured passes. Device event compute time excludes readback; native call includes it. "
                         "No Unity texture upload, display, or Editor FPS. Initialization and BVH build are separate.",
               "results": results}
        (output / "results.json").write_text(json.dumps(doc, indent=2) + "\n")

    try:
        for size in sizes.split(","):
            width, height = map(int, size.split("x"))
            camera.aspect = width / height
            for b in map(int, bounces.split(",")):
                settings.maxBounces = b
                begin = time.perf_counter()
                check(lib.RtInitialize(session, triangle_pointer(triangles), len(triangles), width, height))
                initialization = (time.perf_counter() - begin) * 1000
                pixels = (c.c_ubyte * (width * height * 4))()
                case = {"width": width, "height": height, "bounces": b, "rays_per_pixel": settings.raysPerPixel,
                        "initialize_ms": initialization, "backends": []}
                nodes, depth, build = c.c_int(), c.c_int(), c.c_float()
                lib.RtBvhInfo(session, c.byref(nodes), c.byref(depth), c.byref(build))
                case["bvh"] = {"nodes": nodes.value, "depth": depth.value, "build_ms": build.value}
                reference = None
                for mode, name in enumerate(BACKEND_NAMES):
                    if backend is not None and mode != backend:
                        continue
                    ms = c.c_float()
                    if mode == 3:
                        if not lib.RtRender(session, c.byref(camera), c.byref(settings), pixels, len(pixels), mode,
                                            c.byref(ms)):
                            error = lib.RtLastError(session).decode()
                            case["backends"].append({"backend": name, "unavailable": error})
                            log(f"{name} at {size}, {b} bounces: not available here ({error})")
                            continue
                    lib.RtReset(session)
                    compute, calls = [], []
                    for i in range(passes + 1):
                        begin = time.perf_counter()
                        check(lib.RtRender(session, c.byref(camera), c.byref(settings), pixels, len(pixels), mode,
                                           c.byref(ms)))
                        elapsed = (time.perf_counter() - begin) * 1000
                        if i:
                            compute.append(ms.value)
                            calls.append(elapsed)
                    image = bytes(pixels)
                    key = f"{width}x{height}-b{b}-backend-{mode}"
                    (output / (key + ".rgba")).write_bytes(image)
                    write_png(output / (key + ".png"), image, width, height)
                    if mode == 0:
                        reference = image
                    errors = channel_errors(reference, image) if reference is not None else None
                    maximum = int(errors.max()) if errors is not None and errors.size else None
                    row = {"backend": name, "compute_median_ms": statistics.median(compute),
                           "call_median_ms": statistics.median(calls), "compute_samples_ms": compute,
                           "call_samples_ms": calls, "cpu_threads": lib.RtCpuThreadCount(session),
                           "samples_per_pixel": lib.RtSampleCount(session), "max_channel_error_vs_cpu": maximum,
                           "raw_image_sha256": hashlib.sha256(image).hexdigest(),
                           "mean_abs_channel_error_vs_cpu": float(errors.mean()) if errors is not None else None,
                           "rmse_vs_cpu": float(np.sqrt(np.mean(errors.astype(np.float64) ** 2)))
                           if errors is not None else None,
                           "channels_over_1_fraction": float((errors > 1).mean()) if errors is not None else None,
                           "embree_build_ms": lib.RtEmbreeBuildMilliseconds(session) if mode == 3 else None}
                    case["backends"].append(row)
                    log(f"{name} at {size}, {b} bounces: compute median {row['compute_median_ms']:.3f} ms, "
                        f"largest channel difference from backend 0: {maximum}")
                results.append(case)
                write_results()
    finally:
        lib.RtDestroy(session)
    return json.loads((output / "results.json").read_text())


def main():
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument("snapshot", type=Path)
    parser.add_argument("output", type=Path)
    parser.add_argument("--library", type=Path, default=Path(os.environ.get("RAYTRACER_LIBRARY", LIBRARY)))
    parser.add_argument("--backend", type=int, choices=(0, 1, 2, 3), help="measure one backend only")
    parser.add_argument("--sizes", default="320x180,1280x720")
    parser.add_argument("--bounces", default="3,8")
    parser.add_argument("--passes", type=int, default=5, help="measured passes per backend, after one warm-up")
    args = parser.parse_args()
    run(args.snapshot, args.output, args.library, args.backend, args.sizes, args.bounces, args.passes,
        log=lambda m: print(m, flush=True))


if __name__ == "__main__":
    main()
