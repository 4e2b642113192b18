"""GPU images through RtRender against the oracle at identical sample indices, and the
reference's end-to-end ABI known answers (Tools/verify_native.py) on backends 0, 1, 2."""
import ctypes as c
import math
import os

import numpy as np
import pytest

import oracle
from helpers import cam_record, image_agreement, settings_record, suzanne, vn_quad
from native_api import Camera, Session, Settings, Vector

pytestmark = pytest.mark.gpu


def oracle_frames(tris, cam, st, w, h, calls):
    sc = oracle.OracleScene(tris)
    total, sample = None, 0
    for _ in range(calls):
        rgba, total, rays = sc.render(cam_record(cam), settings_record(st), w, h, sample, total)
        sample += st.raysPerPixel
    return rgba, total, rays


def hdr_mean_relative_error(gpu, ref):
    """Mean over pixels of |gpu - ref| / max(|ref|, 1e-2) on the accumulated radiance."""
    g, r = gpu[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    return float(np.mean(np.abs(g - r) / np.maximum(np.abs(r), 1e-2)))


@pytest.mark.parametrize("calls", [1, 4])
def test_suzanne_gpu_matches_oracle_and_golden(lib, calls):
    tris, cam, st, w, h, golden = suzanne()
    st.seed, st.raysPerPixel = 42, 16
    with Session(lib) as s:
        s.initialize(tris, w, h)
        for _ in range(calls):
            img, _ = s.render(cam, st, 2)
        acc = s.accumulation()
    ref_img, ref_acc, _ = oracle_frames(tris, cam, st, w, h, calls)
    assert hdr_mean_relative_error(acc, ref_acc) <= 1e-3
    mae, over1, _ = image_agreement(img, ref_img)
    assert mae <= 0.5 and over1 <= 0.01
    assert np.array_equal(acc[..., 9], ref_acc[..., 9])          # covered counts
    mae, over1, _ = image_agreement(img, golden["sobol16" if calls == 1 else "sobol64"])
    assert mae <= 0.5 and over1 <= 0.01


def test_cornell_gpu_matches_oracle(lib):
    from scenes import cornell, cornell_camera, make_settings
    w = h = 128
    tris, cam, st = cornell(), cornell_camera(w, h), make_settings(4, rpp=16)
    with Session(lib) as s:
        s.initialize(tris, w, h)
        img, _ = s.render(cam, st, 2)
        acc = s.accumulation()
        closest, shadow = s.ray_counts()
    ref_img, ref_acc, (rc, rs) = oracle_frames(tris, cam, st, w, h, 1)
    assert hdr_mean_relative_error(acc, ref_acc) <= 1e-3
    assert image_agreement(img, ref_img)[0] <= 0.5
    assert abs(closest - rc) <= 1e-3 * rc and abs(shadow - rs) <= 1e-3 * rs


def _glossy_cornell(smoothness):
    from scenes import cornell, srgb_encode
    tris = cornell()
    boxes = tris[12:]
    boxes["specularProbability"] = 1.0
    boxes["smoothness"] = smoothness
    boxes["specularColor"]["x"], boxes["specularColor"]["y"], boxes["specularColor"]["z"] = srgb_encode((0.9, 0.8, 0.6))
    tris[12:] = boxes
    return tris


@pytest.mark.parametrize("case", ["orthographic", "thin_lens", "sky_exposure", "glossy", "mirror", "deep"])
def test_cornell_variants_match_oracle(lib, case):
    """Camera models, sky and exposure, glossy and mirror materials and a long path limit,
    each at identical sample indices against the oracle (Native/renderer.cu:181-288)."""
    from scenes import cornell, make_camera, make_settings
    w, h = 96, 72
    tris = cornell()
    cam = make_camera((0, 1, -3.4), (0, 1, 1), fov_deg=40, aspect=w / h)
    st = make_settings(4, rpp=8)
    if case == "orthographic":
        cam = make_camera((0, 1, -3.4), (0, 1, 1), aspect=w / h, orthographic=1, ortho_size=1.2)
    elif case == "thin_lens":
        st = make_settings(4, rpp=8, aperture=0.08, focus=3.0)
    elif case == "sky_exposure":
        st = make_settings(5, rpp=8, sky=0.6, exposure=2.5)
    elif case == "glossy":
        tris = _glossy_cornell(0.7)
    elif case == "mirror":
        tris = _glossy_cornell(1.0)
    elif case == "deep":
        st = make_settings(24, rpp=4, seed=99)
    with Session(lib) as s:
        s.initialize(tris, w, h)
        img, _ = s.render(cam, st, 2)
        acc = s.accumulation()
        closest, shadow = s.ray_counts()
    ref_img, ref_acc, (rc, rs) = oracle_frames(tris, cam, st, w, h, 1)
    assert hdr_mean_relative_error(acc, ref_acc) <= 1e-3
    mae, over1, _ = image_agreement(img, ref_img)
    assert mae <= 0.5 and over1 <= 0.01
    assert abs(closest - rc) <= 1e-3 * rc and abs(shadow - rs) <= 1e-3 * max(rs, 1)


# ---------------------------------------------------------------- Tools/verify_native.py
LIT = dict(maxBounces=8, exposure=1, aperture=0, focusDistance=3, skyStrength=0, seed=42, raysPerPixel=1)


def lit(**kw):
    d = dict(LIT, **kw)
    return Settings(d["maxBounces"], d["exposure"], d["aperture"], d["focusDistance"], d["skyStrength"], d["seed"],
                    d["raysPerPixel"])


class Agree:
    """verify_native.py:73-95: backends 0, 1 and 2 agree within 1/255; backend 3 (Embree)
    is reported unavailable instead of rendering."""

    def __init__(self, lib):
        self.lib = lib
        self.s = Session(lib)

    def batch(self, cam, st, backend, samples):
        self.lib.RtReset(self.s.handle)
        for _ in range(samples):
            img, ms = self.s.render(cam, st, backend)
            assert ms >= 0 and math.isfinite(ms)
        assert self.lib.RtSampleCount(self.s.handle) == samples * st.raysPerPixel
        return img

    def __call__(self, cam, st, samples=4):
        cpu = self.batch(cam, st, 0, samples)
        assert self.lib.RtCpuThreadCount(self.s.handle) == 1
        multi = self.batch(cam, st, 1, samples)
        if len(os.sched_getaffinity(0)) > 1:
            assert self.lib.RtCpuThreadCount(self.s.handle) > 1
        gpu = self.batch(cam, st, 2, samples)
        assert self.lib.RtCpuThreadCount(self.s.handle) == 0
        assert np.abs(cpu.astype(int) - multi).max() <= 1
        assert np.abs(cpu.astype(int) - gpu).max() <= 1
        return gpu


def centre(img):
    h, w = img.shape[:2]
    return tuple(int(x) for x in img[h // 2, w // 2])


def test_verify_native_known_answers(lib):
    agree = Agree(lib)
    s = agree.s
    try:
        for w, h in [(37, 25), (17, 19)]:
            cam = Camera(Vector(0, 0, -2), Vector(1, 0, 0), Vector(0, 1, 0), Vector(0, 0, 1), .5, w / h, 1, 0)
            for sided, inside, expected in [(0, False, (208, 0, 0, 255)), (1, False, (0, 0, 0, 255)),
                                            (1, True, (208, 0, 0, 255))]:
                red = vn_quad(-1, 1, -1, 1, 0, color=(1, 0, 0), emission=1)
                red["oneSided"] = sided
                cam.origin.z = 2 if inside else -2
                cam.forward.z = -1 if inside else 1
                s.initialize(red, w, h)
                img = agree(cam, lit())
                assert centre(img) == expected
                assert (img[..., 3] == 255).all()
        # independent sessions (verify_native.py:135-143)
        s.initialize(red, 17, 19)
        reference, _ = s.render(cam, lit())
        with Session(lib) as other:
            other.initialize(vn_quad(-1, 1, -1, 1, 0, color=(0, 1, 0), emission=1), 7, 9)
            ocam = Camera(Vector(0, 0, -2), Vector(1, 0, 0), Vector(0, 1, 0), Vector(0, 0, 1), .5, 7 / 9, 1, 0)
            assert centre(other.render(ocam, lit())[0]) == (0, 208, 0, 255)
            lib.RtShutdown(other.handle)
        assert np.array_equal(s.render(cam, lit())[0], reference)

        w, h = 33, 25
        cam = Camera(Vector(0, 0, -3), Vector(1, 0, 0), Vector(0, 1, 0), Vector(0, 0, 1), .3, w / h, 1, 0)
        floor = vn_quad(-2, 2, -2, 2, 0)
        light = vn_quad(.8, 1.2, -.2, .2, -2, emission=30)
        blocker = vn_quad(.25, .75, -.5, .5, -1)
        s.initialize(np.concatenate([floor, light]), w, h)
        unblocked = agree(cam, lit(), 16)
        s.initialize(np.concatenate([floor, light, blocker]), w, h)
        blocked = agree(cam, lit(), 16)
        assert centre(unblocked)[0] > centre(blocked)[0] + 20

        def accumulate(samples, st):
            lib.RtReset(s.handle)
            for _ in range(samples):
                img, _ = s.render(cam, st)
            return img
        rm = lambda a, b: math.sqrt(np.mean((a[..., :3].astype(float) - b[..., :3]) ** 2))
        one, sixteen, ref = accumulate(1, lit()), accumulate(16, lit()), accumulate(256, lit())
        assert rm(sixteen, ref) < rm(one, ref) * .6

        mirror = vn_quad(-2, 2, -2, 2, 0)
        mirror["specularProbability"], mirror["smoothness"] = 1, 1
        behind = vn_quad(-1, 1, -1, 1, -4, color=(0, 1, 0), emission=2)
        s.initialize(np.concatenate([mirror, behind]), w, h)
        img = agree(cam, lit())
        assert centre(img)[1] > 230 and centre(img)[0] == 0

        s.initialize(floor, w, h)
        assert centre(agree(cam, lit())) == (0, 0, 0, 255)
        sky = lit(skyStrength=1)
        assert centre(agree(cam, sky, 16))[0] > 100
        cam.origin.x = 20
        empty = agree(cam, sky)
        assert (empty[..., :3] == 0).all() and (empty[..., 3] == 255).all()
        cam.origin.x = 0
        separate = accumulate(12, sky)
        for rays in (2, 4, 6):
            batched = agree(cam, lit(skyStrength=1, raysPerPixel=rays), 12 // rays)
            assert np.abs(separate.astype(int) - batched).max() <= 1
            assert lib.RtSampleCount(s.handle) == 12
        cam.orthographic = 1
        agree(cam, sky)
        cam.orthographic = 0
        agree(cam, lit(skyStrength=1, aperture=.2))

        # reset / auto-reset (verify_native.py:191-205)
        st = lit(skyStrength=1, aperture=.2)
        lib.RtReset(s.handle)
        a, _ = s.render(cam, st)
        s.render(cam, st)
        assert lib.RtSampleCount(s.handle) == 2
        lib.RtReset(s.handle)
        assert np.array_equal(s.render(cam, st)[0], a) and lib.RtSampleCount(s.handle) == 1
        cam.origin.x += .1
        s.render(cam, st)
        assert lib.RtSampleCount(s.handle) == 1
        st.exposure = 2
        s.render(cam, st)
        assert lib.RtSampleCount(s.handle) == 1
        s.render(cam, st, 1)
        assert lib.RtSampleCount(s.handle) == 1

        # invalid arguments (verify_native.py:206-213)
        out = (c.c_ubyte * (w * h * 4))()
        for backend in (-1, 4):
            assert not lib.RtRender(s.handle, c.byref(cam), c.byref(st), out, len(out), backend, c.byref(c.c_float()))
        assert not lib.RtRender(s.handle, c.byref(cam), c.byref(st), out, 4, 2, c.byref(c.c_float()))
        assert not lib.RtRender(s.handle, c.byref(cam), c.byref(st), out, len(out), 3, c.byref(c.c_float()))
        assert b"Embree" in lib.RtLastError(s.handle)
        s.render(cam, st)
        lib.RtShutdown(s.handle)
        assert not lib.RtRender(s.handle, c.byref(cam), c.byref(st), out, len(out), 2, c.byref(c.c_float()))
    finally:
        s.close()


def test_resident_render_matches_rtrender(lib):
    from scenes import cornell, cornell_camera, make_settings
    tris, cam, st = cornell(), cornell_camera(64, 48), make_settings(4, rpp=4)
    with Session(lib) as a, Session(lib) as b:
        a.initialize(tris, 64, 48)
        b.initialize(tris, 64, 48)
        img, _ = a.render(cam, st)
        ms = b.render_resident(cam, st)
        assert ms > 0 and np.array_equal(b.copy_output(), img)
        assert lib.RtSampleCount(b.handle) == 4


@pytest.mark.parametrize("width", [4, 8])
def test_quantized_nodes_leave_images_unchanged(lib, width):
    """Quantized child boxes only enlarge boxes, so every path makes the same hits: the
    accumulation, the image and the ray counts equal those of full-precision nodes bit for
    bit, on the GPU and on the CPU backend (which traverses the same node array)."""
    from scenes import displaced_icosphere, icosphere_camera, make_settings
    tris = displaced_icosphere(subdiv=5, seed=3)
    cam, st = icosphere_camera(80, 60), make_settings(8, rpp=2, seed=5)
    out = {}
    for q in (0, 1):
        with Session(lib) as s:
            s.set_option("bvh_width", width)
            s.set_option("bvh_quantized", q)
            s.initialize(tris, 80, 60)
            assert lib.RtGetOption(s.handle, b"bvh_quantized") == q
            for _ in range(2):
                img, _ = s.render(cam, st, 2)
            acc = s.accumulation()
            counts = s.ray_counts()
            cpu, _ = s.render(cam, st, 1)
            out[q] = (img, acc, counts, cpu)
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1].view(np.uint32), out[1][1].view(np.uint32))
    assert out[0][2] == out[1][2]
    assert np.array_equal(out[0][3], out[1][3])


def test_frame_launches_pair_shadow_rays_with_the_next_bounce(lib):
    """Each traversal launch of a frame traces one bounce's closest-hit rays together with
    the previous bounce's shadow rays; only the last bounce's shadow rays get a launch of
    their own, and a scene without emitters has none."""
    from scenes import cornell, cornell_camera, make_settings, triangle_soup
    for tris, st, shadow_launches in ((cornell(), make_settings(4, rpp=1), 1),
                                      (triangle_soup(2000, seed=4), make_settings(3, rpp=1, sky=1.0), 0)):
        with Session(lib) as s:
            s.initialize(tris, 64, 48)
            s.set_option("stage_timing", 1)
            s.render_resident(cornell_camera(64, 48), st)
            stages = s.stage_times()
            busy = s.stage_busy("closest", "shadow")
        assert stages["closest"][1] == st.maxBounces
        assert stages["shade"][1] == st.maxBounces
        assert stages["shadow"][1] == shadow_launches
        # One pipeline (a small frame): the launches do not overlap, so their busy time is
        # the sum of their durations.
        total = stages["closest"][0] + stages["shadow"][0]
        assert 0 < busy and abs(busy - total) <= 1e-3 + 1e-3 * total


def test_frames_identical_across_traversal_loops(lib):
    """Closest-hit and shadow rays share launches, and lanes refill across the two queues:
    every loop variant must still make the same hits, so frames agree bit for bit."""
    from scenes import displaced_icosphere, icosphere_camera, make_settings
    tris = displaced_icosphere(subdiv=5, seed=11)
    cam, st = icosphere_camera(96, 64), make_settings(8, rpp=2, seed=9)
    out = []
    for ww, refill, below, stack in ((0, 0, 40, 8), (1, 1, 40, 8), (1, 0, 40, 8), (0, 1, 8, 8), (1, 1, 60, 8),
                                     (1, 1, 40, 12), (0, 0, 40, 16)):
        with Session(lib) as s:
            s.set_option("while_while", ww)
            s.set_option("refill", refill)
            s.set_option("refill_below", below)
            assert s.set_option("short_stack", stack)
            s.set_option("blocks_per_cu", 1)   # few lanes: every lane refills many times
            s.initialize(tris, 96, 64)
            for _ in range(2):
                s.render_resident(cam, st)
            out.append((s.accumulation(), s.ray_counts()))
    for acc, counts in out[1:]:
        assert np.array_equal(acc.view(np.uint32), out[0][0].view(np.uint32))
        assert counts == out[0][1]


@pytest.mark.parametrize("rpp", [1, 3])
def test_path_pipelines_match_one(lib, rpp):
    """Frames split over concurrent path pipelines (own streams, queues and counters; two at
    1 spp here, up to four at 3 spp) equal single-pipeline frames bit for bit, with the same
    ray counts."""
    from scenes import displaced_icosphere, icosphere_camera, make_settings
    tris = displaced_icosphere(subdiv=5, seed=2)
    w, h = 512, 256   # 131072 paths per sample: the size from which a frame splits
    cam, st = icosphere_camera(w, h), make_settings(8, rpp=rpp, seed=4)
    out = {}
    for pipes in (1, 2, 4):
        with Session(lib) as s:
            assert s.set_option("pipelines", pipes)
            s.initialize(tris, w, h)
            for _ in range(2):
                img, _ = s.render(cam, st, 2)
            out[pipes] = (img, s.accumulation(), s.ray_counts())
    for pipes in (2, 4):
        assert np.array_equal(out[1][0], out[pipes][0])
        assert np.array_equal(out[1][1].view(np.uint32), out[pipes][1].view(np.uint32))
        assert out[1][2] == out[pipes][2]


@pytest.mark.parametrize("loop,pipes", [((1, 1), 1), ((0, 0), 2)])
def test_xcd_segment_cursors_match_one_cursor(lib, loop, pipes):
    """Traversal launches that hand out rays from eight segment cursors (one per XCD group of
    workgroups, with stealing once a segment is drained) give the same frames bit for bit and
    the same ray counts as one shared cursor, for both traversal loops, one and two pipelines
    and the path-completion launch.  One workgroup per CU keeps the grid small enough that
    the early bounces take the segmented path (two or more rays per lane) and the late ones
    the shared cursor."""
    from scenes import displaced_icosphere, icosphere_camera, make_settings
    tris = displaced_icosphere(subdiv=5, seed=5)
    w, h = 512, 256
    cam, st = icosphere_camera(w, h), make_settings(8, rpp=4, seed=7)
    out = {}
    for seg in (0, 1):
        with Session(lib) as s:
            assert s.set_option("xcd_segments", seg) and s.get_option("xcd_segments") == seg
            assert s.set_option("while_while", loop[0]) and s.set_option("refill", loop[1])
            assert s.set_option("pipelines", pipes) and s.set_option("tail_bounce", 4)
            assert s.set_option("blocks_per_cu", 1)
            s.initialize(tris, w, h)
            for _ in range(2):
                img, _ = s.render(cam, st, 2)
            out[seg] = (img, s.accumulation(), s.ray_counts())
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1].view(np.uint32), out[1][1].view(np.uint32))
    assert out[0][2] == out[1][2]


@pytest.mark.parametrize("scene", ["icosphere", "cornell_glossy", "soup"])
def test_path_completion_matches_per_bounce_launches(lib, scene):
    """Paths completed in one launch from bounce k on (tail_bounce = k) equal the per-bounce
    launches bit for bit, with the same ray counts, for every k, one or two pipelines."""
    from scenes import (cornell_camera, displaced_icosphere, icosphere_camera, make_settings, soup_camera,
                        triangle_soup)
    w, h = 512, 256
    if scene == "icosphere":
        tris, cam, st = displaced_icosphere(subdiv=5, seed=6), icosphere_camera(w, h), make_settings(8, rpp=1, seed=3)
    elif scene == "cornell_glossy":
        tris, cam, st = _glossy_cornell(0.8), cornell_camera(w, h), make_settings(6, rpp=2, seed=5)
    else:
        tris, cam, st = triangle_soup(30000, seed=2), soup_camera(w, h), make_settings(3, rpp=1, sky=1.0, seed=8)
    out = {}
    for tail in range(0, st.maxBounces + 1):
        for pipes in ((1, 2) if tail in (0, 2) else (2,)):
            with Session(lib) as s:
                assert s.set_option("tail_bounce", tail) and s.set_option("pipelines", pipes)
                s.initialize(tris, w, h)
                for _ in range(2):
                    img, _ = s.render(cam, st, 2)
                out[tail, pipes] = (img, s.accumulation(), s.ray_counts())
    ref = out[0, 1]
    for key, (img, acc, counts) in out.items():
        assert np.array_equal(img, ref[0]), key
        assert np.array_equal(acc.view(np.uint32), ref[1].view(np.uint32)), key
        assert counts == ref[2], key
