"""Unity scene snapshots: the JSON (optionally gzip) files the Unity editor exports.

Format (UnityRayTracer/Assets/RayTracing/RayTraceSession.cs:170-183, written by JsonUtility):
``{"triangles": [Triangle...], "camera": CameraData, "settings": Settings, "width": w,
"height": h}`` with every Vector3 as ``{"x":..,"y":..,"z":..}`` and the struct field names of
the plugin ABI.  Readers of the reference accept a plain or gzip-compressed file
(Tools/verify_quality.py:31-33) and convert each float through FP32 exactly once
(Tools/native_api.py:45-48), which ``load_snapshot`` does too: every JSON number is parsed to
the nearest double and then rounded to the nearest float32, as a ctypes c_float assignment
does, so the triangle bytes handed to RtInitialize are the ones the reference hands over.
"""
import gzip
import hashlib
import json
from pathlib import Path

import numpy as np

from native_api import TRIANGLE_DTYPE, Camera, Settings, Vector, unpack

_VECTOR_FIELDS = ("origin", "edge1", "edge2", "color", "normal0", "normal1", "normal2", "specularColor")
_SCALAR_FIELDS = ("emission", "oneSided", "mesh", "smoothness", "specularProbability")


class Snapshot:
    """One exported scene: triangles (TRIANGLE_DTYPE array), camera, settings and image size."""

    def __init__(self, triangles, camera, settings, width, height, sha256=None):
        self.triangles = triangles
        self.camera = camera
        self.settings = settings
        self.width = int(width)
        self.height = int(height)
        self.sha256 = sha256

    def __repr__(self):
        return f"Snapshot({len(self.triangles)} triangles, {self.width}x{self.height}, sha256={self.sha256})"


def triangles_from_json(items):
    """List of Triangle JSON objects -> contiguous TRIANGLE_DTYPE array."""
    n = len(items)
    out = np.zeros(n, dtype=TRIANGLE_DTYPE)
    for name in _VECTOR_FIELDS:
        vec = np.array([[t[name]["x"], t[name]["y"], t[name]["z"]] for t in items], dtype=np.float64).reshape(n, 3)
        vec = vec.astype(np.float32)
        out[name]["x"], out[name]["y"], out[name]["z"] = vec[:, 0], vec[:, 1], vec[:, 2]
    for name in _SCALAR_FIELDS:
        kind = out.dtype[name]
        out[name] = np.array([t[name] for t in items], dtype=np.float64 if kind.kind == "f" else np.int64).astype(kind)
    return out


def triangles_to_json(triangles):
    """TRIANGLE_DTYPE array -> list of JSON-ready dicts.  Floats are written as the shortest
    decimal that round-trips the float32 value, so load(save(x)) returns the same bytes."""
    items = []
    for t in triangles:
        d = {}
        for name in TRIANGLE_DTYPE.names:
            v = t[name]
            if name in _VECTOR_FIELDS:
                d[name] = {"x": _short(v["x"]), "y": _short(v["y"]), "z": _short(v["z"])}
            elif TRIANGLE_DTYPE[name].kind == "f":
                d[name] = _short(v)
            else:
                d[name] = int(v)
        items.append(d)
    return items


def _short(v):
    """Shortest decimal text of a float32 that parses back (via double) to the same float32."""
    v = np.float32(v)
    for digits in range(6, 10):
        s = float(f"{float(v):.{digits}g}")
        if np.float32(s) == v:
            return s
    return float(v)


def _struct_to_json(obj):
    out = {}
    for name, field in obj._fields_:
        v = getattr(obj, name)
        if field is Vector:
            out[name] = {"x": _short(v.x), "y": _short(v.y), "z": _short(v.z)}
        elif isinstance(v, float):
            out[name] = _short(v)
        else:
            out[name] = int(v)
    return out


def load_snapshot(path):
    """Read a snapshot file (plain or gzip JSON).  ``sha256`` is of the file as stored, the
    identity Tools/benchmark_native.py records for a run."""
    raw = Path(path).read_bytes()
    scene = json.loads(gzip.decompress(raw) if raw[:2] == b"\x1f\x8b" else raw)
    return Snapshot(triangles_from_json(scene["triangles"]), unpack(Camera, scene["camera"]),
                    unpack(Settings, scene["settings"]), scene["width"], scene["height"],
                    hashlib.sha256(raw).hexdigest())


def save_snapshot(path, triangles, camera, settings, width, height, compress=None):
    """Write a snapshot in the editor's format; gzip when ``compress`` (default: name ends in .gz)."""
    path = Path(path)
    doc = {"triangles": triangles_to_json(np.asarray(triangles, dtype=TRIANGLE_DTYPE)),
           "camera": _struct_to_json(camera), "settings": _struct_to_json(settings),
           "width": int(width), "height": int(height)}
    text = json.dumps(doc, separators=(",", ":")).encode()
    if compress is None:
        compress = path.suffix == ".gz"
    path.write_bytes(gzip.compress(text, mtime=0) if compress else text)
    return hashlib.sha256(path.read_bytes()).hexdigest()
