"""Train the learned denoiser on images this path tracer renders from seeded scenes.

Dataset: for each seeded scene (displaced icosphere in a lit room or triangle soup under sky
light, the C3 and C2 generators; optionally Cornell-box variants), a 4-spp render supplies the
network input planes
(RtReadDenoiserInputs: radiance, albedo, normal, depth) and an independently seeded
high-sample render supplies the target radiance.  With --fine, part of the scenes are fine
views instead: full-tessellation icosphere rooms seen through the field of view that a
crop of a 1080-line frame spans, so surface detail covers as many pixels as it does in a
full-resolution frame.  Training runs entirely on the GPU
(RtDenoiserTrain: random crops + flips, Adam, L1 in log space, optionally plus an L1 term on
the display-transformed radiance).  The trained FP32 weights are
written with a versioned header to denoiser_weights.bin next to the library, where RtDenoise
finds them.

    python train_denoiser.py --scenes 96 --target-spp 1024 --steps 3000 --out denoiser_weights.bin

A long run can be split into shards that each render their own scenes (--first-seed) and
continue from the previous shard's weights (--init-weights), with --lr-from / --lr-to placing
each shard on one cosine learning-rate schedule over the whole run.
"""
import argparse
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
if str(HERE) not in sys.path:
    sys.path.insert(0, str(HERE))

import scenes  # noqa: E402
from native_api import Denoiser, Session, load_library  # noqa: E402


def display(hdr, exposure=1.0):
    """Native/renderer.cu:86-90 in float64: exposure tone map, sRGB encode, 8-bit."""
    c = 1 - np.exp(-np.maximum(hdr, 0) * exposure)
    c = np.where(c <= 0.0031308, 12.92 * c, 1.055 * np.power(c, 1 / 2.4) - 0.055)
    return (255 * np.minimum(c, 1) + 0.5).astype(np.uint8)


def psnr(a, b):
    """PSNR of 8-bit display RGB (Tools/verify_quality.py:53-54 RMSE, as 20 log10(255/RMSE))."""
    d = a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)
    rmse = math.sqrt(float(np.mean(d * d)))
    return 99.0 if rmse == 0 else 20 * math.log10(255 / rmse)


def rays_per_call(spp):
    """Samples per RtRender call for an spp-sample image: the largest divisor of spp that is
    at most 16.  Every call must use the same raysPerPixel, since a call whose settings
    differ from the previous one's starts the accumulation over (Native/renderer.cu:388-389)."""
    return max(d for d in range(1, 17) if spp % d == 0)


def render_mean(session, camera, settings, spp, seed):
    """Mean HDR radiance [h][w][3] of spp samples with the given seed."""
    lib = session.lib
    lib.RtReset(session.handle)
    settings.seed = seed
    settings.raysPerPixel = rays_per_call(spp)
    for _ in range(spp // settings.raysPerPixel):
        session.render_resident(camera, settings)
    assert lib.RtSampleCount(session.handle) == spp
    acc = session.accumulation()
    return acc[..., 0:3] / spp


KINDS = ("sphere", "soup", "cornell")


def random_scene(seed, kinds=("sphere", "soup")):
    """(triangles, camera factory, settings) of one seeded training scene: the kind cycles
    with the seed over `kinds` (sphere: displaced icosphere room, soup: triangle soup,
    cornell: Cornell-box variant)."""
    rng = np.random.default_rng(seed)
    kind = KINDS.index(kinds[seed % len(kinds)])
    if kind == 0:
        tris = scenes.displaced_icosphere(subdiv=int(rng.integers(5, 9)), seed=int(seed), amplitude=float(rng.uniform(0.02, 0.1)),
                                          metal_probability=float(rng.uniform(0, 0.5)))
        tris["emission"] = np.where(tris["emission"] > 0, rng.uniform(4, 20), 0).astype(np.float32)
        ang = rng.uniform(-0.8, 0.8)
        dist = rng.uniform(2.4, 3.8)
        eye = (math.sin(ang) * dist, rng.uniform(0.9, 2.4), -math.cos(ang) * dist)
        cam = lambda w, h: scenes.make_camera(eye, (0, 1.25, 0), fov_deg=float(rng.uniform(35, 60)), aspect=w / h)
        st = scenes.make_settings(8, sky=0.0)
    elif kind == 1:
        tris = scenes.triangle_soup(int(rng.integers(20000, 120000)), seed=int(seed))
        eye = (rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(-1.4, -0.8))
        cam = lambda w, h: scenes.make_camera(eye, (0.5, 0.5, 0.5), fov_deg=float(rng.uniform(40, 65)), aspect=w / h)
        st = scenes.make_settings(3, sky=float(rng.uniform(0.5, 2.0)))
    else:
        tris = scenes.cornell()
        wall = rng.uniform(0.1, 0.9, size=3)
        tris["color"]["x"][6:8], tris["color"]["y"][6:8], tris["color"]["z"][6:8] = scenes.srgb_encode(wall)
        tris["emission"][10:12] = rng.uniform(8, 25)
        box = tris[12:].copy()
        for f in ("origin",):
            for a in "xz":
                box[f][a] += rng.uniform(-0.15, 0.15)
        tris = np.concatenate([tris[:12], box])
        eye = (rng.uniform(-0.4, 0.4), rng.uniform(0.7, 1.3), -3.4)
        cam = lambda w, h: scenes.make_camera(eye, (0, 1, 1), fov_deg=40, aspect=w / h)
        st = scenes.make_settings(int(rng.integers(4, 9)))
    return tris, cam, st


def fine_pool(count, subdiv, log=None, first=0, geodesic=False):
    """Displaced-sphere rooms at full tessellation (other seeds than the benchmark scene),
    shared by the fine views: mesh k has seed 5000 + first + k and is the subdivision-`subdiv`
    icosphere, or with `geodesic` every other mesh (odd first + k) the 7,009,280-triangle
    frequency-592 geodesic sphere."""
    pool = []
    for k in range(count):
        seed = 5000 + first + k
        rng = np.random.default_rng(seed)
        amp, metal = float(rng.uniform(0.03, 0.08)), float(rng.uniform(0.1, 0.45))
        if geodesic and (first + k) % 2 == 1:
            tris = scenes.displaced_geodesic(frequency=592, seed=seed, amplitude=amp, metal_probability=metal)
        else:
            tris = scenes.displaced_icosphere(subdiv=subdiv, seed=seed, amplitude=amp, metal_probability=metal)
        tris["emission"] = np.where(tris["emission"] > 0, rng.uniform(8, 20), 0).astype(np.float32)
        pool.append(tris)
        if log:
            log(f"fine mesh {k + 1}/{count}: {len(tris)} triangles")
    return pool


def fine_view(seed, size, full_height=1080):
    """A size x size window of a full_height-line frame: the camera of a displaced-icosphere
    room aimed at a random point of the sphere, with the field of view a window of that
    size spans in the full frame, so features cover as many pixels as at full resolution."""
    rng = np.random.default_rng(seed)
    ang, dist = rng.uniform(-0.8, 0.8), rng.uniform(2.8, 4.2)
    eye = (math.sin(ang) * dist, rng.uniform(0.9, 2.4), -math.cos(ang) * dist)
    d = rng.normal(size=3)
    target = np.array([0, 1.3, 0]) + d / np.linalg.norm(d) * rng.uniform(0, 1.1)
    full = math.radians(rng.uniform(35, 60))
    fov = 2 * math.degrees(math.atan(math.tan(full / 2) * size / full_height))
    return scenes.make_camera(eye, tuple(target), fov_deg=fov, aspect=1.0), scenes.make_settings(8, sky=0.0)


def heldout_scene(seed):
    """A 256 x 256 view of a seeded scene of any of the three kinds (for tests)."""
    tris, cam, st = random_scene(seed, KINDS)
    return cam(256, 256), st, tris


def build_dataset(lib, count, size, input_spp, target_spp, first_seed, log, fine=0.0, pool=(), kinds=("sphere", "soup")):
    """input_spp: one sample count, or a sequence cycled over the scenes (mixed noise levels).
    A fraction `fine` of the scenes (chosen by seed) are fine views of the meshes in pool."""
    spps = tuple(input_spp) if isinstance(input_spp, (tuple, list)) else (input_spp,)
    inputs = np.zeros((count, size, size, 10), np.float32)
    targets = np.zeros((count, size, size, 3), np.float32)
    sessions = {}
    t0 = time.perf_counter()
    try:
        for i in range(count):
            seed = first_seed + i
            if pool and np.random.default_rng(seed + 777).uniform() < fine:
                k = i % len(pool)
                if k not in sessions:
                    sessions[k] = Session(lib)
                    sessions[k].initialize(pool[k], size, size)
                s, own = sessions[k], False
                cam, st = fine_view(seed, size)
            else:
                tris, camf, st = random_scene(seed, kinds)
                cam = camf(size, size)
                s, own = Session(lib), True
                s.initialize(tris, size, size)
            try:
                targets[i] = render_mean(s, cam, st, target_spp, seed=1000003 + i)
                render_mean(s, cam, st, spps[i % len(spps)], seed=42 + i)
                inputs[i] = s.denoiser_inputs()
            finally:
                if own:
                    s.close()
            if log and (i + 1) % 16 == 0:
                log(f"dataset {i + 1}/{count} ({time.perf_counter() - t0:.1f} s)")
    finally:
        for s in sessions.values():
            s.close()
    return inputs, targets


def evaluate(net, inputs, targets):
    """Mean display-space PSNR of plain 4-spp and of the network, vs the targets."""
    out = net.infer(inputs)
    p_in = np.mean([psnr(display(inputs[i, ..., 0:3]), display(targets[i])) for i in range(len(inputs))])
    p_out = np.mean([psnr(display(out[i]), display(targets[i])) for i in range(len(inputs))])
    return float(p_in), float(p_out)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--scenes", type=int, default=96)
    ap.add_argument("--heldout", type=int, default=8)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--input-spp", type=int, default=4)
    ap.add_argument("--input-spp-mix", default=None,
                    help="comma-separated input sample counts cycled over the training scenes, e.g. 4,4,8,16,32 "
                         "(held-out scenes stay at --input-spp)")
    ap.add_argument("--target-spp", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--display-weight", type=float, default=0.0,
                    help="weight of the display-space L1 term added to the log-space L1 loss")
    ap.add_argument("--fine", type=float, default=0.0,
                    help="fraction of scenes that are fine views: full-tessellation icosphere rooms seen "
                         "through the field of view of a crop of a 1080-line frame")
    ap.add_argument("--fine-meshes", type=int, default=4)
    ap.add_argument("--fine-subdiv", type=int, default=9)
    ap.add_argument("--fine-first", type=int, default=0, help="seed offset of the first fine mesh")
    ap.add_argument("--fine-geodesic", action="store_true",
                    help="every other fine mesh is the frequency-592 geodesic sphere instead of the icosphere")
    ap.add_argument("--kinds", default="sphere,soup",
                    help="scene generators cycled over the seeds: sphere (C3), soup (C2), cornell")
    ap.add_argument("--first-seed", type=int, default=100, help="seed of the first training scene")
    ap.add_argument("--init-weights", default=None, help="continue from these weights instead of a fresh network")
    ap.add_argument("--lr-from", type=float, default=0.0,
                    help="where this run starts on the cosine learning-rate schedule (0 = lr, 1 = 0)")
    ap.add_argument("--lr-to", type=float, default=1.0, help="where this run ends on that schedule")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=str(HERE / "denoiser_weights.bin"))
    ap.add_argument("--report", default=None, help="write a JSON summary here")
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    lib = load_library()
    t0 = time.perf_counter()
    mix = tuple(int(x) for x in args.input_spp_mix.split(",")) if args.input_spp_mix else args.input_spp
    kinds = tuple(k for k in args.kinds.split(",") if k)
    if not kinds or any(k not in KINDS for k in kinds):
        raise SystemExit(f"--kinds: choose from {KINDS}")
    pool = fine_pool(args.fine_meshes, args.fine_subdiv, log, args.fine_first, args.fine_geodesic) if args.fine > 0 else ()
    inputs, targets = build_dataset(lib, args.scenes, args.size, mix, args.target_spp, args.first_seed, log, args.fine,
                                    pool, kinds)
    hin, htg = build_dataset(lib, args.heldout, args.size, args.input_spp, args.target_spp, 100000, log, args.fine,
                             pool, kinds)
    del pool
    data_s = time.perf_counter() - t0
    net = Denoiser(lib, weights=args.init_weights) if args.init_weights else Denoiser(lib, seed=args.seed)
    net.set_display_loss(args.display_weight)
    net.set_dataset(inputs, targets)
    before = evaluate(net, hin, htg)
    log(f"held-out PSNR before training: plain {before[0]:.2f} dB, network {before[1]:.2f} dB")
    losses = []
    t0 = time.perf_counter()
    chunks = 10
    for k in range(chunks):
        at = args.lr_from + (args.lr_to - args.lr_from) * k / chunks   # cosine decay in chunks
        lr = args.lr * 0.5 * (1 + math.cos(math.pi * at))
        losses.append(net.train(args.steps // chunks, args.batch, args.crop, lr, args.seed + k))
        log(f"steps {(k + 1) * (args.steps // chunks)}: lr {lr:.2e}, loss {losses[-1][-50:].mean():.4f}")
    train_s = time.perf_counter() - t0
    fwd_ms, step_ms = net.timings()
    after = evaluate(net, hin, htg)
    log(f"held-out PSNR after training: plain {after[0]:.2f} dB, network {after[1]:.2f} dB")
    net.save(args.out)
    flops = 3 * lib.RtDenoiserForwardFlops(args.batch, args.crop, args.crop)
    report = {"scenes": args.scenes, "first_seed": args.first_seed, "kinds": args.kinds,
              "init_weights": args.init_weights, "lr_from": args.lr_from, "lr_to": args.lr_to,
              "heldout": args.heldout, "size": args.size, "input_spp": args.input_spp,
              "input_spp_mix": args.input_spp_mix, "fine": args.fine, "fine_meshes": args.fine_meshes,
              "fine_subdiv": args.fine_subdiv, "fine_first": args.fine_first, "fine_geodesic": args.fine_geodesic,
              "target_spp": args.target_spp, "steps": args.steps, "batch": args.batch, "crop": args.crop,
              "display_weight": args.display_weight,
              "dataset_seconds": round(data_s, 1), "train_seconds": round(train_s, 1),
              "train_step_ms": round(step_ms, 3),
              "train_step_mfma_utilization": round(flops / (step_ms / 1e3) / 2.5e15, 4) if step_ms > 0 else None,
              "loss_first": float(np.concatenate(losses)[:20].mean()), "loss_last": float(np.concatenate(losses)[-100:].mean()),
              "heldout_psnr_plain_db": round(after[0], 2), "heldout_psnr_network_db": round(after[1], 2),
              "heldout_psnr_untrained_db": round(before[1], 2), "weights": str(args.out)}
    print(json.dumps(report), flush=True)
    if args.report:
        Path(args.report).write_text(json.dumps(report, indent=1))
    net.close()


if __name__ == "__main__":
    main()
