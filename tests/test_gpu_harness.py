"""The Unity call sequence (compiled C, RayTraceSession.cs:135-238) and the benchmark
harness (Tools/benchmark_native.py + compare_benchmarks.py) run against the GPU build."""
import json
import subprocess

import numpy as np
import pytest

import scenes
from benchmark_native import run as benchmark
from compare_benchmarks import compare
from helpers import image_agreement, suzanne
from native_api import LIBRARY, Camera, Session, Settings
from snapshot import save_snapshot
from test_harness_cpu import compile_unity_harness

pytestmark = pytest.mark.gpu


def test_unity_call_sequence(lib, tmp_path):
    """UploadScene -> Calculate x4 -> Render(denoise) -> camera move -> backend switches ->
    Dispose, through dlsym'd exports and C-declared pack-4 structs; the first GPU frame is
    byte-identical to the same call made through the Python binding."""
    exe = compile_unity_harness(tmp_path)
    tris = scenes.cornell()
    (tmp_path / "tris.bin").write_bytes(tris.tobytes())
    w, h = 96, 64
    print(f"unity harness: cornell {len(tris)} tris {w}x{h}", flush=True)
    out = tmp_path / "frame.rgba"
    run = subprocess.run([str(exe), str(LIBRARY), "session", str(tmp_path / "tris.bin"), str(len(tris)), str(w), str(h),
                          str(out)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr + run.stdout
    report = json.loads(run.stdout)
    assert report["failures"] == 0 and report["denoised"] == 1, report
    assert report["samples_after_frames"] == 8 and report["embree"] == 0
    data = out.read_bytes()
    frame = np.frombuffer(data[:w * h * 4], np.uint8).reshape(h, w, 4)
    cam = Camera.from_buffer_copy(data[w * h * 4:w * h * 4 + 64])
    st = Settings.from_buffer_copy(data[w * h * 4 + 64:])
    with Session(lib) as s:
        s.initialize(tris, w, h)
        img, _ = s.render(cam, st, 2)
    assert np.array_equal(img, frame)
    assert frame[..., :3].max() > 0


def test_benchmark_runs_are_hash_identical(tmp_path):
    """Two benchmark runs of one snapshot pass the exact-hash gate on every backend; CPU
    single and CPU multi agree bit for bit and the GPU image passes the reference's
    CPU-vs-CUDA image gate (Tools/verify_demo_scenes.py:54-56)."""
    tris, cam, st, w, h, _ = suzanne()
    st.raysPerPixel, st.seed = 2, 42
    snap = tmp_path / "suzanne.json.gz"
    save_snapshot(snap, tris, cam, st, w, h)
    docs = []
    for name in ("before", "after"):
        print(f"benchmark run {name}", flush=True)
        docs.append(benchmark(snap, tmp_path / name, LIBRARY, sizes="64x36,96x54", bounces="3,8", passes=2,
                              log=lambda m: print(m, flush=True)))
    comparison = compare(*docs)
    assert len(comparison["results"]) == 4 * 3 and all(r["image_identical"] for r in comparison["results"])
    for case in docs[1]["results"]:
        rows = {r["backend"]: r for r in case["backends"]}
        assert "unavailable" in rows["Embree CPU multi"]
        assert rows["CPU single"]["raw_image_sha256"] == rows["CPU multi"]["raw_image_sha256"]
        assert rows["CPU single"]["cpu_threads"] == 1 and rows["HIP GPU"]["cpu_threads"] == 0
        assert rows["HIP GPU"]["samples_per_pixel"] == 3 * st.raysPerPixel
        gpu = rows["HIP GPU"]
        assert gpu["mean_abs_channel_error_vs_cpu"] <= 0.5 and gpu["channels_over_1_fraction"] <= 0.01, gpu
        key = f'{case["width"]}x{case["height"]}-b{case["bounces"]}'
        a = np.frombuffer((tmp_path / "after" / f"{key}-backend-0.rgba").read_bytes(), np.uint8)
        b = np.frombuffer((tmp_path / "after" / f"{key}-backend-2.rgba").read_bytes(), np.uint8)
        shape = (case["height"], case["width"], 4)
        assert image_agreement(a.reshape(shape), b.reshape(shape))[0] == pytest.approx(
            gpu["mean_abs_channel_error_vs_cpu"])


EMBREE_CHECK = r"""
import json, sys
sys.path[:0] = sys.argv[1:4]
import numpy as np
import scenes
from helpers import image_agreement
from native_api import Session, load_library
lib = load_library()
tris = scenes.cornell()
w, h = 48, 36
cam, st = scenes.cornell_camera(w, h), scenes.make_settings(4, rpp=4)
with Session(lib) as s:
    s.initialize(tris, w, h)
    cpu, _ = s.render(cam, st, 1)
    emb, _ = s.render(cam, st, 3)
    out = {"build_ms": lib.RtEmbreeBuildMilliseconds(s.handle), "threads": lib.RtCpuThreadCount(s.handle),
           "samples": lib.RtSampleCount(s.handle), "agreement": image_agreement(cpu, emb),
           "lit": int(emb[..., :3].max())}
print(json.dumps(out))
"""


def test_embree_backend_through_an_embree_abi_library(tmp_path):
    """Backend 3 end to end: dlopen of libembree3.so.3 (here a brute-force test double with the
    Embree 3 ABI slice backend 3 calls, tests/embree_double.c), vertex/index upload, the
    determinant filter and the result mapping.  Its image matches the CPU backend's within
    the reference's CPU-vs-GPU image gate (Embree receives vertices, not edges, so t, u, v can
    differ by an ulp)."""
    import os
    import shutil
    import sys
    from pathlib import Path
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    here = Path(__file__).resolve().parent
    subprocess.run([cc, "-O2", "-shared", "-fPIC", "-Wall", "-Werror", str(here / "embree_double.c"), "-o",
                    str(tmp_path / "libembree3.so.3"), "-lm"], check=True)
    env = dict(os.environ, LD_LIBRARY_PATH=f"{tmp_path}:{os.environ.get('LD_LIBRARY_PATH', '')}")
    pkg = str(here.parent / "custom-ray-tracing-shader_amd")
    run = subprocess.run([sys.executable, "-c", EMBREE_CHECK, pkg, str(here), str(here.parent)], capture_output=True, text=True,
                         timeout=120, env=env)
    assert run.returncode == 0, run.stderr[-2000:]
    r = json.loads(run.stdout.strip().splitlines()[-1])
    assert r["build_ms"] > 0 and r["threads"] >= 1 and r["samples"] == 4 and r["lit"] > 0, r
    mae, over1, _ = r["agreement"]
    assert mae <= 0.5 and over1 <= 0.01, r
