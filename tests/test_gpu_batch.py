"""Batched views: every view of a batch equals a session rendering that scene alone with
one sample per call (same seed and sample indices), bit for bit, with and without the
captured graph; sample counters advance and reset."""
import numpy as np
import pytest

import scenes
from native_api import Batch, Session

pytestmark = pytest.mark.gpu


def views():
    out = []
    t = scenes.cornell()
    out.append((t, 64, 48, scenes.cornell_camera(64, 48), scenes.make_settings(4, seed=7)))
    t = scenes.triangle_soup(20000, seed=5)
    out.append((t, 80, 60, scenes.soup_camera(80, 60), scenes.make_settings(3, sky=1.0, seed=42)))
    t = scenes.displaced_icosphere(subdiv=4, seed=3)
    out.append((t, 96, 64, scenes.icosphere_camera(96, 64), scenes.make_settings(8, seed=1234)))
    t = scenes.cornell()
    out.append((t, 33, 17, scenes.cornell_camera(33, 17), scenes.make_settings(2, seed=9, aperture=0.05, focus=3.0)))
    return out


@pytest.mark.parametrize("use_graph,width,quantized", [(1, 4, 0), (0, 4, 0), (1, 8, 0), (0, 8, 0), (1, 4, 1), (1, 8, 1)])
def test_batch_views_equal_single_sessions(lib, use_graph, width, quantized):
    vs = views()
    frames = 3
    with Batch(lib, use_graph=use_graph, bvh_width=width, bvh_quantized=quantized) as b:
        for t, w, h, cam, st in vs:
            b.add_view(t, w, h, cam, st)
        b.build()
        b.render(frames)
        closest, shadow = b.ray_counts()
        got = [b.read(v) for v in range(len(vs))]
    expect_closest = expect_shadow = 0
    for v, (t, w, h, cam, st) in enumerate(vs):
        with Session(lib) as s:
            s.set_option("bvh_width", width)
            s.set_option("bvh_quantized", quantized)
            s.initialize(t, w, h)
            st.raysPerPixel = 1
            for _ in range(frames):
                img, _ = s.render(cam, st)
            acc = s.accumulation()
            c, sh = s.ray_counts()
            expect_closest += c
            expect_shadow += sh
        assert np.array_equal(got[v][0], img), v
        assert np.array_equal(got[v][1], acc), v
    assert (closest, shadow) == (expect_closest, expect_shadow)


def test_batch_reset_and_counts(lib):
    vs = views()[:2]
    with Batch(lib) as b:
        for t, w, h, cam, st in vs:
            b.add_view(t, w, h, cam, st)
        b.build()
        b.render(1)
        first = b.read(0)[0]
        b.render(2)
        import ctypes as c
        n = c.c_uint()
        assert lib.RtBatchViewInfo(b.handle, 1, None, None, c.byref(n)) and n.value == 3
        b.reset()
        b.render(1)
        assert np.array_equal(b.read(0)[0], first)
        assert not lib.RtBatchAddView(b.handle, None, 0, 1, 1, None, None)


@pytest.mark.parametrize("use_graph", [1, 0])
def test_batch_pipelines_match_one(lib, use_graph):
    """A batch frame split over concurrent path pipelines (captured across streams in the
    graph) equals the single-pipeline frame bit for bit."""
    vs = views()
    t = scenes.displaced_icosphere(subdiv=4, seed=8)
    vs.append((t, 512, 512, scenes.icosphere_camera(512, 512), scenes.make_settings(6, seed=77)))  # 262K+ slots
    out = {}
    for pipes in (1, 2, 4):
        with Batch(lib, use_graph=use_graph, pipelines=pipes) as b:
            for tv, w, h, cam, st in vs:
                b.add_view(tv, w, h, cam, st)
            b.build()
            b.render(2)
            out[pipes] = ([b.read(v) for v in range(len(vs))], b.ray_counts())
    for pipes in (2, 4):
        assert out[pipes][1] == out[1][1]
        for v in range(len(vs)):
            assert np.array_equal(out[pipes][0][v][0], out[1][0][v][0]), (pipes, v)
            assert np.array_equal(out[pipes][0][v][1].view(np.uint32), out[1][0][v][1].view(np.uint32)), (pipes, v)


def test_batch_xcd_segments_match_one_cursor(lib):
    """Batch launches that hand out rays from eight segment cursors (one workgroup per CU, so
    the early bounces take the segmented path) equal the one-cursor frames bit for bit, with
    the graph and two pipelines."""
    vs = views()
    t = scenes.displaced_icosphere(subdiv=4, seed=8)
    vs.append((t, 512, 512, scenes.icosphere_camera(512, 512), scenes.make_settings(6, seed=77)))
    out = {}
    for seg in (0, 1):
        with Batch(lib, pipelines=2, blocks_per_cu=1, xcd_segments=seg) as b:
            for tv, w, h, cam, st in vs:
                b.add_view(tv, w, h, cam, st)
            b.build()
            b.render(2)
            out[seg] = ([b.read(v) for v in range(len(vs))], b.ray_counts())
    assert out[0][1] == out[1][1]
    for v in range(len(vs)):
        assert np.array_equal(out[0][0][v][0], out[1][0][v][0]), v
        assert np.array_equal(out[0][0][v][1].view(np.uint32), out[1][0][v][1].view(np.uint32)), v
