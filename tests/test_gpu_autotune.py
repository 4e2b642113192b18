"""The library's launch autotuner (RtAutotune, RtBatchAutotune): searches on small scenes
return configurations the library accepts and that render the defaults' images bit for bit,
persist in the cache, and reach fresh sessions and batches of the same class through
RtInitialize / RtBatchBuild, except for the options their caller set."""
import json

import numpy as np
import pytest

import autotune
import scenes
from native_api import Batch, Session

pytestmark = pytest.mark.gpu


def test_session_autotune_reaches_fresh_sessions(lib, tmp_path, monkeypatch):
    cache = tmp_path / "cache.json"
    monkeypatch.setenv("RT_AUTOTUNE_CACHE", str(cache))
    tris = scenes.displaced_icosphere(subdiv=4, seed=2)
    w, h = 96, 64
    cam, st = scenes.icosphere_camera(w, h), scenes.make_settings(8, seed=3)
    best, trials = autotune.autotune(tris, cam, st, w, h, frames=2, fixed={"bvh_width": 4}, lib=lib)
    assert trials > 50 and best["mrays_per_s"] > 0 and best["bvh_width"] == 4
    entry = json.loads(cache.read_text())[f"tris2^{round(np.log2(len(tris)))}-px{w}x{h}-b8"]
    assert {k: entry[k] for k in autotune.KEYS} == {k: best[k] for k in autotune.KEYS}
    # a fresh session of the scene class takes the cached launch at RtInitialize ...
    with Session(lib) as s:
        s.initialize(tris, w, h)
        assert {k: s.get_option(k) for k in autotune.KEYS} == {k: best[k] for k in autotune.KEYS}
        tuned_img, _ = s.render(cam, st)
    # ... except for the options its caller set, and nothing of it with autotune_cache 0
    other = 2 if best["blocks_per_cu"] != 2 else 3
    with Session(lib) as s:
        assert s.set_option("blocks_per_cu", other)
        s.initialize(tris, w, h)
        s.render(cam, st)
        assert s.get_option("blocks_per_cu") == other and s.get_option("short_stack") == best["short_stack"]
    with Session(lib) as s:
        assert s.set_option("autotune_cache", 0)
        s.initialize(tris, w, h)
        default_img, _ = s.render(cam, st)
        assert s.get_option("blocks_per_cu") == 3 and s.get_option("build_max_leaf") == 4
    assert np.array_equal(tuned_img, default_img)


def test_batch_autotune_reaches_fresh_batches(lib, tmp_path, monkeypatch):
    monkeypatch.setenv("RT_AUTOTUNE_CACHE", str(tmp_path / "cache.json"))
    vs = [(scenes.cornell(), 48, 32, scenes.cornell_camera(48, 32), scenes.make_settings(4, seed=7)),
          (scenes.triangle_soup(5000, seed=5), 40, 30, scenes.soup_camera(40, 30), scenes.make_settings(3, sky=1.0))]
    best, trials = autotune.autotune_batch(vs, frames=2, fixed={"bvh_width": 4}, lib=lib)
    assert trials > 20 and best["mrays_per_s"] > 0
    out = []
    for options in ({}, {"autotune_cache": 0}):
        with Batch(lib, **options) as b:
            for t, w, h, cam, st in vs:
                b.add_view(t, w, h, cam, st)
            b.build()
            got = {k: lib.RtBatchGetOption(b.handle, k.encode()) for k in autotune.BATCH_KEYS}
            if not options:
                assert got == {k: best[k] for k in autotune.BATCH_KEYS}
            b.render(2)
            out.append([b.read(v) for v in range(len(vs))])
    for a, c in zip(out[0], out[1]):
        assert np.array_equal(a[0], c[0]) and np.array_equal(a[1].view(np.uint32), c[1].view(np.uint32))
