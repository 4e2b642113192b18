"""Write the golden fixtures used by the tests from the reference's own data files.

Reads (as data, never executes) the reference's Suzanne scene snapshot and its golden CUDA
renders of the current Sobol build:
  Benchmarks/embree/scene.json               982 triangles, camera, settings
  Benchmarks/quality/suzanne/sobol-16.png    seed 42, 16 spp, 320x180 (Tools/verify_quality.py:35-52)
  Benchmarks/quality/suzanne/sobol-64.png    seed 42, 64 spp
  Benchmarks/quality/suzanne/reference.png   seed 20261023, 2048 spp
and stores them in suzanne.npz (images flipped back to the ABI's bottom-row-first order,
since verify_quality.py saves with FLIP_TOP_BOTTOM).  Run: python tests/golden/make_fixtures.py
"""
import json
import sys
from pathlib import Path

import numpy as np
from PIL import Image

REF = Path("/root/reference")
HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1] / "custom-ray-tracing-shader_amd"))
from native_api import TRIANGLE_DTYPE  # noqa: E402


def scene_arrays(scene):
    tris = np.zeros(len(scene["triangles"]), dtype=TRIANGLE_DTYPE)
    for name in TRIANGLE_DTYPE.names:
        col = [t[name] for t in scene["triangles"]]
        if isinstance(col[0], dict):
            for a in "xyz":
                tris[name][a] = [v[a] for v in col]
        else:
            tris[name] = col
    cam = scene["camera"]
    camera = np.array([cam[k][a] for k in ("origin", "right", "up", "forward") for a in "xyz"]
                      + [cam["tanHalfFov"], cam["aspect"], cam["orthographicSize"]], dtype=np.float32)
    st = scene["settings"]
    settings = np.array([st["maxBounces"], st["exposure"], st["aperture"], st["focusDistance"], st["skyStrength"],
                         st["seed"], st["raysPerPixel"]], dtype=np.float64)
    return tris, camera, np.int32(cam["orthographic"]), settings


def png(path):
    return np.asarray(Image.open(path).convert("RGBA"))[::-1].copy()


def main():
    scene = json.loads((REF / "Benchmarks/embree/scene.json").read_text())
    tris, camera, ortho, settings = scene_arrays(scene)
    q = REF / "Benchmarks/quality/suzanne"
    np.savez_compressed(HERE / "suzanne.npz", triangles=tris.view(np.uint8), camera=camera, orthographic=ortho,
                        settings=settings, sobol16=png(q / "sobol-16.png"), sobol64=png(q / "sobol-64.png"),
                        reference2048=png(q / "reference.png"), width=np.int32(320), height=np.int32(180))
    print("wrote", HERE / "suzanne.npz")


if __name__ == "__main__":
    main()
