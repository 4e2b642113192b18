"""The launch autotuner: searches run on small scenes, return a configuration the library
accepts (and that renders the same image as the defaults), and persist in the cache."""
import numpy as np
import pytest

import autotune
import scenes
from native_api import Batch, Session

pytestmark = pytest.mark.gpu

KEYS = ("bvh_width", "blocks_per_cu", "short_stack", "while_while", "refill", "pipelines")


def test_session_autotune_returns_valid_cached_config(lib, tmp_path, monkeypatch):
    monkeypatch.setenv("RT_AUTOTUNE_CACHE", str(tmp_path / "cache.json"))
    tris = scenes.displaced_icosphere(subdiv=4, seed=2)
    w, h = 96, 64
    cam, st = scenes.icosphere_camera(w, h), scenes.make_settings(8, seed=3)
    best, trials = autotune.autotune(tris, cam, st, w, h, frames=2, widths=(4,), lib=lib)
    assert trials and all(t["mrays_per_s"] > 0 for t in trials)
    assert best["mrays_per_s"] == max(t["mrays_per_s"] for t in trials)
    for k in KEYS + ("build_max_leaf", "tail_bounce"):
        assert k in best
    # the second search of the same scene class comes from the cache
    again, none = autotune.autotune(tris, cam, st, w, h, lib=lib)
    assert none == [] and again == best
    # the chosen launch renders the image the defaults render
    imgs = []
    for cfg in ({}, best):
        with Session(lib) as s:
            s.set_option("build_max_leaf", best["build_max_leaf"])
            if cfg:
                autotune.apply(s, cfg)
            s.initialize(tris, w, h)
            img, _ = s.render(cam, st)
            imgs.append(img)
    assert np.array_equal(imgs[0], imgs[1])


def test_batch_autotune_returns_working_config(lib):
    vs = [(scenes.cornell(), 48, 32, scenes.cornell_camera(48, 32), scenes.make_settings(4, seed=7)),
          (scenes.triangle_soup(5000, seed=5), 40, 30, scenes.soup_camera(40, 30), scenes.make_settings(3, sky=1.0))]
    best, trials = autotune.autotune_batch(vs, widths=(4,), lib=lib)
    assert trials and best["mrays_per_s"] == max(t["mrays_per_s"] for t in trials)
    cfg = {k: best[k] for k in KEYS}
    out = []
    for options in ({}, cfg):
        with Batch(lib, **options) as b:
            for t, w, h, cam, st in vs:
                b.add_view(t, w, h, cam, st)
            b.build()
            b.render(1)
            out.append([b.read(v)[0] for v in range(len(vs))])
    for a, c in zip(out[0], out[1]):
        assert np.array_equal(a, c)
