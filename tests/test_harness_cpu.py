"""Host-side harnesses around the plugin: Unity scene snapshots, the benchmark comparison
gate (Tools/compare_benchmarks.py) and the Unity P/Invoke binding (RayTraceSession.cs)
compiled as C.  None of these tests launches device work."""
import copy
import json
import shutil
import struct
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

import scenes
from benchmark_native import channel_errors, write_png
from compare_benchmarks import ComparisonError, compare
from helpers import GOLDEN, suzanne
from native_api import LIBRARY, TRIANGLE_DTYPE
from snapshot import load_snapshot, save_snapshot

HERE = Path(__file__).resolve().parent
REFERENCE = Path("/root/reference")


# ------------------------------------------------------------------ snapshots
@pytest.mark.parametrize("name", ["scene.json", "scene.json.gz"])
def test_snapshot_round_trip_is_bit_exact(tmp_path, name):
    tris, cam, st, w, h, _ = suzanne()
    path = tmp_path / name
    digest = save_snapshot(path, tris, cam, st, w, h)
    snap = load_snapshot(path)
    assert snap.triangles.dtype == TRIANGLE_DTYPE
    assert snap.triangles.tobytes() == tris.tobytes()
    assert bytes(snap.camera) == bytes(cam) and bytes(snap.settings) == bytes(st)
    assert (snap.width, snap.height) == (w, h)
    assert snap.sha256 == digest
    assert (path.read_bytes()[:2] == b"\x1f\x8b") == name.endswith(".gz")


def test_snapshot_round_trip_of_generated_scene(tmp_path):
    tris = scenes.displaced_icosphere(subdiv=3, seed=11)
    cam = scenes.icosphere_camera(64, 36)
    st = scenes.make_settings(5, rpp=3, seed=7, sky=0.25)
    save_snapshot(tmp_path / "ico.json", tris, cam, st, 64, 36)
    snap = load_snapshot(tmp_path / "ico.json")
    assert snap.triangles.tobytes() == tris.tobytes()
    assert bytes(snap.camera) == bytes(cam) and bytes(snap.settings) == bytes(st)


def test_snapshot_fields_follow_the_editor_format(tmp_path):
    tris, cam, st, w, h, _ = suzanne()
    save_snapshot(tmp_path / "s.json", tris[:2], cam, st, w, h)
    doc = json.loads((tmp_path / "s.json").read_text())
    assert set(doc) == {"triangles", "camera", "settings", "width", "height"}
    assert set(doc["triangles"][0]) == set(TRIANGLE_DTYPE.names)
    assert set(doc["triangles"][0]["origin"]) == {"x", "y", "z"}
    assert set(doc["camera"]) == {"origin", "right", "up", "forward", "tanHalfFov", "aspect", "orthographicSize",
                                  "orthographic"}
    assert set(doc["settings"]) == {"maxBounces", "exposure", "aperture", "focusDistance", "skyStrength", "seed",
                                    "raysPerPixel"}


@pytest.mark.skipif(not REFERENCE.exists(), reason="reference checkout not present")
def test_reference_snapshots_load():
    """The editor-written Suzanne snapshot gives the triangles of the committed fixture, and the
    gzip Dragon snapshot loads with its recorded triangle count (87,165)."""
    snap = load_snapshot(REFERENCE / "Benchmarks/embree/scene.json")
    tris, cam, *_ = suzanne()
    assert snap.sha256.startswith("2a7a4cc6")
    assert snap.triangles.tobytes() == tris.tobytes()
    assert bytes(snap.camera)[:48] == bytes(cam)[:48]
    dragon = load_snapshot(REFERENCE / "Benchmarks/quality/dragon/scene.json.gz")
    assert len(dragon.triangles) == 87165 and dragon.settings.maxBounces == 8


# ------------------------------------------------------------------ compare gate
def _runs():
    d = GOLDEN / "bvh_benchmark"
    return (json.loads((d / "before.json").read_text()), json.loads((d / "after.json").read_text()),
            json.loads((d / "comparison.json").read_text()))


def test_compare_reproduces_reference_comparison():
    """Benchmarks/bvh/{before,after}.json through the gate give Benchmarks/bvh/comparison.json."""
    before, after, expected = _runs()
    assert compare(before, after) == expected


@pytest.mark.parametrize("mutate,message", [
    (lambda b, a: a["results"][1]["backends"][2].__setitem__("raw_image_sha256", "0" * 64), "output hash of backend CUDA GPU changed"),
    (lambda b, a: a.__setitem__("snapshot_sha256", "f" * 64), "different scene snapshots"),
    (lambda b, a: a.__setitem__("measured_passes", 6), "different number of passes"),
    (lambda b, a: a.__setitem__("warmup_passes", 2), "different number of warm-up passes"),
    (lambda b, a: a.__setitem__("omp_threads", "8"), "different OpenMP thread settings"),
    (lambda b, a: a["results"][0].__setitem__("bounces", 4), "cover different"),
    (lambda b, a: a["results"][0]["backends"][0].__setitem__("samples_per_pixel", 7), "rendered a different workload"),
    (lambda b, a: a["results"][0]["backends"][1].__setitem__("cpu_threads", 4), "rendered a different workload"),
])
def test_compare_rejects_every_mismatch(mutate, message):
    before, after, _ = _runs()
    after = copy.deepcopy(after)
    mutate(before, after)
    with pytest.raises(ComparisonError, match=message):
        compare(before, after)


def test_compare_carries_unavailable_backends():
    before, after, expected = _runs()
    for doc in (before, after):
        for case in doc["results"]:
            case["backends"].append({"backend": "Embree CPU multi", "unavailable": "not installed"})
    assert compare(before, after) == expected
    after["results"][0]["backends"][-1] = dict(before["results"][0]["backends"][0], backend="Embree CPU multi")
    with pytest.raises(ComparisonError, match="available in one run only"):
        compare(before, after)


# ------------------------------------------------------------------ image output
def _decode_png(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        assert struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0] == zlib.crc32(kind + body) & 0xFFFFFFFF
        chunks.setdefault(kind, b"")
        chunks[kind] += body
        pos += 12 + n
    w, h, depth, colour = struct.unpack(">IIBB", chunks[b"IHDR"][:10])
    raw = zlib.decompress(chunks[b"IDAT"])
    rows = [raw[y * (4 * w + 1):(y + 1) * (4 * w + 1)] for y in range(h)]
    assert all(r[0] == 0 for r in rows) and (depth, colour) == (8, 6)
    return np.frombuffer(b"".join(r[1:] for r in rows), np.uint8).reshape(h, w, 4)


def test_png_is_top_row_first(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (5, 7, 4), dtype=np.uint8)        # bottom row first, as RtRender writes
    write_png(tmp_path / "x.png", img.tobytes(), 7, 5)
    assert np.array_equal(_decode_png((tmp_path / "x.png").read_bytes()), img[::-1])


def test_channel_errors_skip_alpha():
    a = np.array([10, 20, 30, 255, 0, 0, 0, 255], np.uint8).tobytes()
    b = np.array([12, 20, 27, 0, 0, 1, 0, 0], np.uint8).tobytes()
    assert channel_errors(a, b).tolist() == [2, 0, 3, 0, 1, 0]


# ------------------------------------------------------------------ Unity binding in C
def compile_unity_harness(out_dir):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = Path(out_dir) / "unity_harness"
    subprocess.run([cc, "-O2", "-std=c11", "-Wall", "-Wextra", "-Werror", str(HERE / "unity_harness.c"), "-o", str(exe),
                    "-ldl", "-lm"], check=True)
    return exe


def test_unity_binding_resolves_exports_and_null_contract(tmp_path):
    """C structs declared as the Unity side does (pack 4, static size/offset asserts), all 13
    exports resolved with dlsym, and the null-session contract of Native/renderer.cu:310-318."""
    exe = compile_unity_harness(tmp_path)
    run = subprocess.run([str(exe), str(LIBRARY), "bind"], capture_output=True, text=True, timeout=60)
    assert run.returncode == 0, run.stderr
    assert json.loads(run.stdout) == {"failures": 0, "exports": 13}
