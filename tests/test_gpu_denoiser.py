"""GPU denoiser kernels and RtDenoise against the FP64 oracle (oracle/denoiser.py) and the
reference's RtDenoise contract (Tools/verify_denoising.py:37-86)."""
import ctypes as c
from pathlib import Path

import numpy as np
import pytest

from oracle import denoiser as odn
from native_api import Camera, Denoiser, Session, Settings, Vector, as_pointer

pytestmark = pytest.mark.gpu
WEIGHTS = Path(__file__).resolve().parents[1] / "custom-ray-tracing-shader_amd" / "denoiser_weights.bin"


def conv_forward(lib, x, w, b, relu):
    n, h, wd, cin = x.shape
    cout = w.shape[1]
    out = np.zeros((n, h, wd, cout), np.float32)
    x, w, b = (np.ascontiguousarray(a, np.float32) for a in (x, w, b))
    assert lib.RtConv3x3Forward(as_pointer(x, c.c_float), n, h, wd, cin, as_pointer(w, c.c_float),
                                as_pointer(b, c.c_float), cout, relu, as_pointer(out, c.c_float))
    return out


# Images at least 16 pixels wide (the first shape and the last nine) run on the LDS-tiled
# kernel, with partial 8 x 64 tiles, one or several input slices and one or several
# output-channel groups; the others on the small-image kernel.  The last four run 8-row tiles
# with every input slice's weights resident in LDS: 32-channel groups over 2 and 4 slices,
# 64-channel groups over 2 slices, and 32-channel groups chosen because 64-channel groups'
# weights (128 inputs) would not fit.
@pytest.mark.parametrize("cin,cout,h,w,relu", [(10, 32, 16, 16, 1), (32, 64, 5, 7, 0), (224, 96, 8, 12, 1),
                                               (32, 3, 9, 4, 0), (64, 128, 6, 6, 1), (160, 64, 4, 8, 0),
                                               (10, 32, 130, 150, 1), (96, 64, 136, 128, 0), (32, 3, 129, 131, 0),
                                               (224, 96, 21, 70, 1), (96, 128, 13, 50, 0), (64, 64, 130, 260, 1),
                                               (128, 64, 130, 200, 0), (64, 64, 256, 512, 1), (128, 64, 256, 512, 0)])
def test_conv_forward_vs_fp64(lib, cin, cout, h, w, relu):
    rng = np.random.default_rng(cin * 1000 + cout)
    x = odn.bf16_round(rng.normal(size=(2, h, w, cin)).astype(np.float32))
    wt = odn.bf16_round((rng.normal(size=(9, cout, cin)) / np.sqrt(9 * cin)).astype(np.float32))
    b = rng.normal(size=cout).astype(np.float32)
    got = conv_forward(lib, x, wt, b, relu)
    ref = odn.conv3x3(x.astype(np.float64), wt.astype(np.float64), b.astype(np.float64))
    if relu:
        ref = np.maximum(ref, 0)
    # bf16 x bf16 products are exact in FP32; only the FP32 accumulation order differs.
    scale = np.abs(x).max() * np.abs(wt).max() * 9 * cin
    assert np.abs(got - ref).max() <= 1e-6 * scale + 1e-6


@pytest.mark.parametrize("cin,cout,h,w", [(32, 32, 16, 16), (224, 96, 8, 8), (64, 128, 12, 10), (32, 3, 16, 16),
                                          (64, 32, 130, 132), (96, 128, 10, 52)])
def test_conv_backward_vs_fp64(lib, cin, cout, h, w):
    rng = np.random.default_rng(cin + cout)
    n = 2
    x = odn.bf16_round(rng.normal(size=(n, h, w, cin)).astype(np.float32))
    wt = odn.bf16_round((rng.normal(size=(9, cout, cin)) / np.sqrt(9 * cin)).astype(np.float32))
    g = odn.bf16_round(rng.normal(size=(n, h, w, cout)).astype(np.float32))
    gi = np.zeros_like(x)
    gw = np.zeros_like(wt)
    gb = np.zeros(cout, np.float32)
    assert lib.RtConv3x3Backward(as_pointer(x, c.c_float), n, h, w, cin, as_pointer(wt, c.c_float),
                                 as_pointer(g, c.c_float), cout, as_pointer(gi, c.c_float), as_pointer(gw, c.c_float),
                                 as_pointer(gb, c.c_float))
    rx, rw, rb = odn.conv3x3_backward(x.astype(np.float64), wt.astype(np.float64), g.astype(np.float64))
    assert np.abs(gi - rx).max() <= 1e-5 * np.abs(rx).max() + 1e-6
    assert np.abs(gw - rw).max() <= 1e-5 * np.abs(rw).max() + 1e-6
    assert np.abs(gb - rb).max() <= 1e-5 * np.abs(rb).max() + 1e-6


@pytest.mark.parametrize("display_weight", [0.0, 4.0])
def test_network_gradients_vs_fp64_oracle(lib, display_weight):
    rng = np.random.default_rng(7)
    raw = np.abs(rng.normal(size=(2, 32, 32, 10))).astype(np.float32)
    raw[..., 6:9] = rng.normal(size=(2, 32, 32, 3))
    target = (np.abs(rng.normal(size=(2, 32, 32, 3))) * 4 + 2).astype(np.float32)
    with Denoiser(lib, seed=11) as net:
        net.set_display_loss(display_weight)
        params = net.params
        loss, grads = net.gradients(raw, target)
    ref_loss, ref_grads = odn.loss_and_grads(odn.bf16_round(params).astype(np.float64), raw, target, display_weight)
    assert loss == pytest.approx(ref_loss, rel=2e-2)
    lay, _ = odn.layout()
    for name, cin, cout, ci, co, relu, w_off, b_off in lay:
        g = grads[w_off:b_off + co].astype(np.float64)
        r = ref_grads[w_off:b_off + co]
        cos = float(g @ r / (np.linalg.norm(g) * np.linalg.norm(r) + 1e-30))
        rel = float(np.linalg.norm(g - r) / (np.linalg.norm(r) + 1e-30))
        # BF16 activations and gradients between layers: a wrong wire or missing term gives
        # cosine ~0 or relative error ~1; rounding gives a few percent.
        assert cos > 0.99 and rel < 0.12, (name, cos, rel)


def test_inference_vs_fp64_oracle(lib):
    rng = np.random.default_rng(9)
    raw = np.abs(rng.normal(size=(1, 48, 32, 10))).astype(np.float32)
    with Denoiser(lib, seed=5) as net:
        params = net.params
        out = net.infer(raw)
    ref = odn.infer(odn.bf16_round(params).astype(np.float64), raw)
    assert np.abs(out - ref).max() <= 0.03 * np.abs(ref).max()


def test_inference_with_upsampling_read_in_place_vs_fp64_oracle(lib):
    """At 512 x 528 the decoder's first convolutions at levels 0-2 run on the tile kernel, which
    reads their upsampled channels from the half-resolution tensor in place (partial 64-wide
    tiles included); the result is the oracle's network."""
    rng = np.random.default_rng(19)
    raw = np.abs(rng.normal(size=(1, 512, 528, 10))).astype(np.float32)
    raw[..., 6:9] = rng.normal(size=(1, 512, 528, 3))
    with Denoiser(lib, seed=7) as net:
        params = net.params
        out = net.infer(raw)
    ref = odn.infer(odn.bf16_round(params).astype(np.float64), raw)
    assert np.abs(out - ref).max() <= 0.03 * np.abs(ref).max()


def test_batched_inference_vs_fp64_oracle(lib):
    """Two different images in one call (RtDenoiserInfer with n = 2): at 144 x 272 the level-0
    layers run the 8-row kernels that load their halos straight into LDS (D0A's upsampled
    channels included) and level 1 the 4-row kernels, so both find image 1's pixels behind
    image 0's."""
    rng = np.random.default_rng(23)
    raw = np.abs(rng.normal(size=(2, 144, 272, 10))).astype(np.float32)
    raw[1] *= 3.0
    raw[..., 6:9] = rng.normal(size=(2, 144, 272, 3))
    with Denoiser(lib, seed=11) as net:
        params = net.params
        out = net.infer(raw)
    ref = odn.infer(odn.bf16_round(params).astype(np.float64), raw)
    for i in range(2):
        assert np.abs(out[i] - ref[i]).max() <= 0.03 * np.abs(ref[i]).max(), i


def test_training_reduces_loss(lib):
    rng = np.random.default_rng(3)
    n, h, w = 6, 64, 64
    yy, xx = np.mgrid[0:h, 0:w] / 16.0
    clean = np.stack([1 + 0.5 * np.sin(xx + k) * np.cos(yy - k) for k in range(3)], -1)[None].repeat(n, 0)
    clean = (clean * rng.uniform(0.5, 2.0, (n, 1, 1, 1))).astype(np.float32)
    noisy = np.maximum(clean * (1 + 0.6 * rng.normal(size=clean.shape)), 0).astype(np.float32)
    raw = np.zeros((n, h, w, 10), np.float32)
    raw[..., 0:3] = noisy
    raw[..., 3:6] = 0.5
    raw[..., 7] = 1
    raw[..., 9] = 3
    with Denoiser(lib, seed=2) as net:
        net.set_dataset(raw, clean)
        losses = net.train(150, batch=4, crop=32, lr=1e-3, seed=1)
        _, step_ms = net.timings()
    assert np.isfinite(losses).all()
    assert losses[-20:].mean() < 0.6 * losses[:5].mean(), (losses[:5], losses[-20:])
    assert step_ms > 0


# ---------------------------------------------------------------- RtDenoise contract
def _scene():
    from scenes import cornell, cornell_camera, make_settings
    return cornell(), cornell_camera, make_settings(4, rpp=4)


def test_rtdenoise_contract(lib):
    tris, camf, st = _scene()
    with Denoiser(lib, seed=3) as net:
        for w, h in ((97, 55), (37, 25)):
            cam = camf(w, h)
            with Session(lib) as s:
                s.initialize(tris, w, h)
                lib.RtSetDenoiser(s.handle, net.handle)
                out = (c.c_ubyte * (w * h * 4))()
                ms = c.c_float()
                assert not lib.RtDenoise(s.handle, out, len(out), c.byref(ms)), "must render before denoising"
                ref = None
                for backend in (0, 1, 2):
                    lib.RtReset(s.handle)
                    s.render(cam, st, backend)
                    before, version = lib.RtSampleCount(s.handle), lib.RtAccumulationVersion(s.handle)
                    img, ms2 = s.denoise()
                    assert lib.RtSampleCount(s.handle) == before and lib.RtAccumulationVersion(s.handle) == version
                    assert ms2 > 0 and (img[..., 3] == 255).all()
                    if ref is None:
                        ref = img.astype(int)
                    assert np.abs(img.astype(int) - ref).max() <= 1, backend
                assert not lib.RtDenoise(s.handle, out, len(out) - 1, c.byref(ms))
                version = lib.RtAccumulationVersion(s.handle)
                cam.origin.x += .01
                s.render(cam, st)
                assert lib.RtSampleCount(s.handle) == 4 and lib.RtAccumulationVersion(s.handle) != version


def test_rtdenoise_preserves_accumulation_and_black_background(lib):
    tris, camf, st = _scene()
    w, h = 37, 25
    cam = camf(w, h)
    with Denoiser(lib, seed=4) as net, Session(lib) as a, Session(lib) as b:
        for s in (a, b):
            s.initialize(tris, w, h)
            lib.RtSetDenoiser(s.handle, net.handle)
        first, _ = a.render(cam, st)
        second, _ = b.render(cam, st)
        assert np.array_equal(first, second)
        a.denoise(); b.denoise(); a.denoise()
        assert np.array_equal(a.render(cam, st)[0], b.render(cam, st)[0])
        lib.RtShutdown(b.handle)
        out = (c.c_ubyte * (w * h * 4))()
        assert not lib.RtDenoise(b.handle, out, len(out), c.byref(c.c_float()))
        cam.origin.x = 10000
        a.render(cam, st)
        img, _ = a.denoise()
        assert (img[..., :3] == 0).all() and (img[..., 3] == 255).all()


def test_forced_fallback_is_plain_accumulation(lib):
    tris, camf, st = _scene()
    w, h = 64, 48
    cam = camf(w, h)
    with Denoiser(lib, seed=5) as net, Session(lib) as s:
        s.initialize(tris, w, h)
        lib.RtSetDenoiser(s.handle, net.handle)
        assert s.set_option("denoise_fallback", 1)
        plain, _ = s.render(cam, st)
        img, _ = s.denoise()
        tiles, rejected, _ = s.denoise_stats()
        assert tiles == 4 * 3 and rejected == tiles
        assert np.array_equal(img, plain)
        assert s.set_option("denoise_fallback", 2)
        s.denoise()
        assert s.denoise_stats()[1] == 0


@pytest.mark.skipif(not WEIGHTS.exists(), reason="trained weights not built (train_denoiser.py)")
def test_trained_network_beats_plain_accumulation(lib):
    """On a held-out scene, the trained network at 4 spp is closer to a 1024-spp render than
    plain 4-spp accumulation (8-bit display PSNR, as Tools/verify_quality.py:53-54)."""
    import train_denoiser as td
    cam, st, tris = td.heldout_scene(12345)
    w = h = 256
    with Session(lib) as s:
        s.initialize(tris, w, h)
        ref = td.render_mean(s, cam, st, 1024, seed=777)
        lib.RtReset(s.handle)
        st.raysPerPixel, st.seed = 4, 42
        noisy, _ = s.render(cam, st)
        den, _ = s.denoise()
    ref_img = td.display(ref, st.exposure)
    gain = td.psnr(den, ref_img) - td.psnr(noisy, ref_img)
    assert gain > 3.0, gain


@pytest.mark.skipif(not WEIGHTS.exists(), reason="trained weights not built (train_denoiser.py)")
def test_denoise_ensemble_passes(lib):
    """denoise_ensemble 2 and 4 add network passes on mirrored inputs (their outputs mirrored
    back and averaged): the frames change, the reported network time grows, and the quality
    stays that of one pass (within 0.5 dB here)."""
    import train_denoiser as td
    cam, st, tris = td.heldout_scene(4321)
    w = h = 128
    out = {}
    with Session(lib) as s:
        s.initialize(tris, w, h)
        ref_img = td.display(td.render_mean(s, cam, st, 512, seed=99), st.exposure)
        lib.RtReset(s.handle)
        st.raysPerPixel, st.seed = 8, 42
        s.render(cam, st)
        for e in (1, 2, 4):
            assert s.set_option("denoise_ensemble", e)
            den, _ = s.denoise()
            out[e] = (den, td.psnr(den, ref_img), s.denoise_stats()[2])
        assert not s.set_option("denoise_ensemble", 3)
    assert not np.array_equal(out[1][0], out[2][0]) and not np.array_equal(out[2][0], out[4][0])
    assert out[4][2] > out[2][2] > out[1][2] > 0
    for e in (2, 4):
        assert out[e][1] > out[1][1] - 0.5, (e, out[e][1], out[1][1])


@pytest.mark.skipif(not WEIGHTS.exists(), reason="trained weights not built (train_denoiser.py)")
def test_variance_check_rejects_corrupted_network(lib):
    """With the per-tile check on (denoise_fallback 0), the trained network passes most tiles of
    a Cornell frame; the same network made about ten times too bright (ln 10 added to its three
    output biases: the output is a residual in log space) fails every tile, and the frame is
    exactly the plain accumulation, as the reference shows the raw image when its denoiser
    cannot run (UnityRayTracer/Assets/RayTracing/RayTraceSession.cs:213-217)."""
    tris, camf, st = _scene()
    w, h = 96, 80
    cam = camf(w, h)
    with Denoiser(lib, weights=WEIGHTS) as net, Session(lib) as s:
        s.initialize(tris, w, h)
        lib.RtSetDenoiser(s.handle, net.handle)
        assert s.get_option("denoise_fallback") == 0
        plain, _ = s.render(cam, st)
        good, _ = s.denoise()
        tiles, rejected, _ = s.denoise_stats()
        assert tiles == 6 * 5 and rejected < tiles // 4, rejected
        assert not np.array_equal(good, plain)
        params = net.params
        params[-32:-29] += np.log(10.0)   # the output layer's biases (32 padded channels, 3 used)
        net.params = params
        bad, _ = s.denoise()
        tiles, rejected, _ = s.denoise_stats()
        assert rejected == tiles, (rejected, tiles)
        assert np.array_equal(bad, plain)
