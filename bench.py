"""Mrays/s of the MI355X path tracer on one GPU, with roofline, CPU-baseline and
time-to-equal-PSNR figures.

Workload (default, "geodesic"): the procedural large scene - a frequency-592 geodesic sphere
(7,009,280 triangles) displaced by 6-octave fractal noise (amplitude .05), Lambertian or
glossy-metal triangles (p = .3, roughness U(0,1)), inside a 20-triangle room with two
emitters - at 1920x1080, 8 bounces + Russian roulette, 1 progressive sample per step, BVH
built once at RtInitialize.  A step is one RtRenderResident call (raygen, traversal, shading
and accumulation of one frame, output left in HBM).  value = (closest-hit + shadow ray
queries) / wall time of the timed steps.  Other workloads: icosphere9 (the 5,242,900-triangle
subdivision-9 icosphere of the same kind), soup (1M random triangles), cornell, batch.

After the timed region, instrumented frames measure (a) per-launch kernel time with HIP
events on each launch's own stream and (b) node visits / triangle tests per ray, from which
the traversal kernel's algorithmic bytes per launch are formed:
    bytes = 48 B per ray (32 B ray in + 16 B hit out) + node visits * node bytes
            + triangle tests * 48 B
roofline.achieved = those bytes per launch / the traversal's busy time per launch (the union
of its launch intervals: launches of concurrent path pipelines overlap).  It is a rate of
algorithmic bytes, not of HBM traffic; roofline.hbm_traffic_frac gives the measured HBM bytes
(rocprof, profiles/traffic_<workload>.json) over the same time.

time_to_equal_psnr is measured live on the same session (C3 workloads): plain progressive
accumulation and the learned denoiser against a 4096-spp render (Tools/verify_quality.py:53-80).

The CPU baseline is the oracle (a plain C++ restatement of the reference's CPU path,
OpenMP over image rows like Native/renderer.cu:405-416) on the host's cores, over a
bounded block of rows of the same frame.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for p in (str(ROOT), str(ROOT / "custom-ray-tracing-shader_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
# Launch configurations chosen by the library's autotuner (RtAutotune / RtBatchAutotune) on an
# MI355X for each workload (bench.py --autotune repeats the search); the BVH width also fixes
# the node size below.  Leaf size
# changes the work per ray, never the hits or images: the soup runs leaf <= 2 (its
# configuration's leaf <= 4 renders identical frames, tests/test_gpu_configs.py).
TUNED = {
    "soup": {"bvh_width": 4, "build_max_leaf": 2, "build_bins": 32, "blocks_per_cu": 2, "short_stack": 12, "while_while": 1, "refill": 1,
             "pipelines": 4},
    "icosphere": {"bvh_width": 4, "build_max_leaf": 2, "blocks_per_cu": 3, "short_stack": 12, "while_while": 1,
                  "refill": 1, "pipelines": 4, "tail_bounce": 4},
    "cornell": {"bvh_width": 4, "build_max_leaf": 4, "blocks_per_cu": 8, "short_stack": 12, "while_while": 1, "refill": 0,
                "pipelines": 4},
    "batch": {"bvh_width": 4, "blocks_per_cu": 3, "short_stack": 12, "while_while": 1, "refill": 1, "pipelines": 3},
}
TUNED["geodesic"] = TUNED["icosphere9"] = TUNED["icosphere"]
C3_WORKLOADS = ("geodesic", "icosphere9", "icosphere")


def bench_batch(args):
    import native_api
    import scenes
    views = scenes.batch_views(args.batch_views)
    cfg = dict(TUNED["batch"])
    for k in ("bvh_width", "bvh_quantized", "blocks_per_cu", "short_stack", "refill", "refill_below", "while_while",
              "pipelines", "xcd_segments"):
        if getattr(args, k) is not None:
            cfg[k] = getattr(args, k)
    if args.autotune:
        import autotune
        t0 = time.perf_counter()
        best, trials = autotune.autotune_batch(views)
        cfg = {k: v for k, v in best.items() if k != "mrays_per_s"}
        log(f"autotune (library search): {trials} batch configurations in {time.perf_counter() - t0:.1f} s, "
            f"best {best}")
    b = native_api.Batch(**cfg)
    for t, w, h, cam, st in views:
        b.add_view(t, w, h, cam, st)
    build_ms = b.build()
    pixels = sum(w * h for _, w, h, _, _ in views)
    tris = sum(len(t) for t, *_ in views)
    log(f"batch: {len(views)} views, {tris} triangles, {pixels} pixels; build {build_ms:.0f} ms")
    b.render(args.warmup)
    rays, wall = 0, 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.render(1)
        c, sh = b.ray_counts()
        rays += c + sh
    wall = time.perf_counter() - t0
    b.close()
    line = {"metric": "Mrays/s on 1 MI355X (batched views)", "value": round(rays / wall / 1e6, 2), "unit": "Mrays/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded scene generators)",
            "config": {"workload": f"{len(views)} batched views (soup / displaced sphere), one persistent launch per "
                                   f"stage, hipGraph per frame", "views": len(views), "triangles": tris,
                       "pixels": pixels, "launch": cfg, "autotuned_now": args.autotune,
                       "build_ms": round(build_ms, 1)}}
    print(json.dumps(line), flush=True)
    return 0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(name, tris_count):
    import scenes
    if name == "soup":
        w, h = 1920, 1080
        tris = scenes.triangle_soup(tris_count, seed=1)
        return (f"random triangle soup {len(tris)} tris, 1920x1080, 3 bounces, 1 spp/step, sky 1, "
                f"GPU binned-SAH build {TUNED['soup']['build_bins']} bins leaf <= {TUNED['soup']['build_max_leaf']}",
                tris, scenes.soup_camera(w, h), scenes.make_settings(3, rpp=1, sky=1.0), w, h)
    if name == "geodesic":
        w, h = 1920, 1080
        tris = scenes.displaced_geodesic(frequency=592, seed=7)
        return (f"displaced frequency-592 geodesic sphere + 20-tri room {len(tris)} tris, 1920x1080, "
                f"8 bounces + RR, 1 spp/step",
                tris, scenes.icosphere_camera(w, h), scenes.make_settings(8, rpp=1, sky=0.0), w, h)
    if name in ("icosphere9", "icosphere"):
        w, h = 1920, 1080
        tris = scenes.displaced_icosphere(subdiv=9, seed=7)
        return (f"displaced subdivision-9 icosphere + 20-tri room {len(tris)} tris, 1920x1080, 8 bounces + RR, "
                f"1 spp/step",
                tris, scenes.icosphere_camera(w, h), scenes.make_settings(8, rpp=1, sky=0.0), w, h)
    if name == "cornell":
        w = h = 256
        return ("Cornell box 36 tris, 256x256, 4 bounces, 16 spp/step", scenes.cornell(), scenes.cornell_camera(w, h),
                scenes.make_settings(4, rpp=16), w, h)
    raise SystemExit(f"unknown workload {name}")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(tris, cam, st, w, h, budget_s):
    """Oracle on all host cores over whole-frame row blocks until about budget_s of work."""
    import oracle
    scene = oracle.OracleScene(tris)
    cam_rec = np.frombuffer(bytes(cam), dtype=oracle.CAMERA).copy()
    st_rec = np.frombuffer(bytes(st), dtype=oracle.SETTINGS).copy()
    threads = len(os.sched_getaffinity(0))
    total = np.zeros((h, w, 10), dtype=np.float32)
    rays, rows, sample, elapsed = 0, 0, 0, 0.0
    block = max(1, h // 16)
    y = 0
    while elapsed < budget_s:
        y1 = min(h, y + block)
        t0 = time.perf_counter()
        _, _, (c, sh) = scene.render(cam_rec, st_rec, w, h, sample, total, threads, rows=(y, y1))
        elapsed += time.perf_counter() - t0
        rays += c + sh
        rows += y1 - y
        y = y1
        if y >= h:
            y, sample = 0, sample + int(st_rec["raysPerPixel"][0])
    return {"value": rays / elapsed / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"oracle (C++ restatement of the reference CPU path, 12-bin SAH BVH, OpenMP over rows, "
                      f"{threads} threads on {cpu_model()}) on {rows} rows x {w} px of the same frame and settings: "
                      f"{rays} ray queries in {elapsed:.1f} s (scene BVH build not timed)"}


def denoise_timing(args):
    """RtDenoise on a 4-spp 1080p frame of a C3 workload, repeated: the workload for profiling
    the denoiser's kernels (rocprofv3 --kernel-trace) apart from the study."""
    import native_api
    name = args.workload if args.workload in C3_WORKLOADS else "geodesic"
    desc, tris, cam, st, w, h = workload(name, 0)
    s = native_api.Session()
    for k, v in TUNED[name].items():
        s.set_option(k, v)
    s.initialize(tris, w, h)
    st.raysPerPixel = 4
    s.render(cam, st)
    times, network = [], []
    for _ in range(args.denoise_timing):
        _, ms = s.denoise()
        times.append(ms)
        network.append(s.denoise_stats()[2])
    tiles, rejected, infer_ms = s.denoise_stats()
    flops = s.lib.RtDenoiserForwardFlops(1, (h + 15) // 16 * 16, (w + 15) // 16 * 16)
    s.close()
    print(json.dumps({"workload": desc + ", 4 spp", "denoise_calls": len(times),
                      "rtdenoise_ms_min": round(min(times), 3), "network_ms_min": round(min(network), 3),
                      "network_ms_median": round(float(np.median(network)), 3),
                      "network_flop": flops, "tiles": tiles, "tiles_fallback": rejected}), flush=True)
    return 0


def denoise_train_timing(args):
    """--denoise-train-timing STEPS: RtDenoiserTrain steps (forward, backward, Adam) on random
    data of --train-batch crops of --train-crop pixels, after 20 warm-up steps: the workload
    for profiling the training kernels.  MFMA utilisation counts 3 x the forward FLOP."""
    import native_api
    lib = native_api.load_library()
    b, c = args.train_batch, args.train_crop
    rng = np.random.default_rng(0)
    n = max(b, 8)
    inputs = rng.uniform(0, 1, (n, c, c, 10)).astype(np.float32)
    targets = rng.uniform(0, 1, (n, c, c, 3)).astype(np.float32)
    with native_api.Denoiser(lib, seed=1) as net:
        net.set_display_loss(4.0)
        net.set_dataset(inputs, targets)
        net.train(20, b, c, 1e-4, 1)
        t0 = time.perf_counter()
        net.train(args.denoise_train_timing, b, c, 1e-4, 2)
        wall = (time.perf_counter() - t0) / args.denoise_train_timing * 1e3
        step_ms = net.timings()[1]
    flops = 3 * lib.RtDenoiserForwardFlops(b, c, c)
    print(json.dumps({"batch": b, "crop": c, "steps": args.denoise_train_timing, "step_ms": round(step_ms, 3),
                      "wall_ms_per_step": round(wall, 3), "train_flop": flops,
                      "mfma_utilization": round(flops / (step_ms / 1e3) / 2.5e15, 4)}), flush=True)
    return 0


def run_psnr_study(s, cam, st, w, h, reference_spp, denoised_ladder=(8, 12, 16, 20, 24, 32, 64), repeats=3,
                    ensemble_ladder=(12, 16, 20)):
    """Time to equal PSNR on an initialized session: learned denoise vs plain accumulation,
    against a reference_spp render (seed 20261023), 8-bit display PSNR as
    Tools/verify_quality.py:53-54.  Wall times include each RtRender's readback
    (Tools/verify_quality.py:72-78) and the denoiser's transfers.  The denoised ladder runs
    with the session's denoise_ensemble, then ensemble_ladder with two network passes
    (denoise_ensemble 2: the input mirrored, the outputs averaged); the 256-spp match is the
    fastest denoised configuration that reaches plain 256-spp PSNR.  Restores the session's
    ensemble setting; its accumulation is left at the last denoised frame."""
    import copy
    import train_denoiser as td
    lib = s.lib
    st = copy.copy(st)
    ensemble = lib.RtGetOption(s.handle, b"denoise_ensemble")
    t0 = time.perf_counter()
    ref = td.render_mean(s, cam, st, reference_spp, seed=20261023)
    ref_s = time.perf_counter() - t0
    ref_img = td.display(ref, st.exposure)
    ref_albedo = s.denoiser_inputs()[..., 3:6]

    def plain(spp, seed=42):
        lib.RtReset(s.handle)
        st.seed = seed
        st.raysPerPixel = td.rays_per_call(spp)   # equal calls: changed settings restart accumulation
        t = time.perf_counter()
        for _ in range(spp // st.raysPerPixel):
            img, _ = s.render(cam, st)
        assert lib.RtSampleCount(s.handle) == spp
        return img, time.perf_counter() - t

    ladder = []
    for spp in (4, 8, 16, 32, 64, 128, 256, 512):
        best = None
        for _ in range(2 if spp <= 256 else 1):
            img, sec = plain(spp)
            best = sec if best is None else min(best, sec)
        ladder.append({"spp": spp, "seconds": round(best, 4), "psnr_db": round(td.psnr(img, ref_img), 3)})
        log(f"plain {spp:4d} spp: {best * 1e3:8.1f} ms, PSNR {ladder[-1]['psnr_db']:.2f} dB")
    plain(4)          # warm the denoiser path once (weights load, buffers)
    s.denoise()

    def denoised(spp):
        best = None
        for _ in range(repeats):
            t = time.perf_counter()
            plain_img, _ = plain(spp)
            img, _ = s.denoise()
            el = time.perf_counter() - t
            if best is None or el < best[0]:
                best = (el, img, plain_img)
        return best

    sec, img, plain_img = denoised(4)
    tiles, rejected, infer_ms = s.denoise_stats()
    dn_psnr = td.psnr(img, ref_img)
    # The ceiling of any denoiser that re-modulates its irradiance estimate with the frame's own
    # 4-sample albedo (as this one does): the reference's exact irradiance times that albedo.
    albedo4 = s.denoiser_inputs()[..., 3:6]
    eps = 0.01   # the network's demodulation offset (denoiser.hip kDemodEps)
    ceiling = td.psnr(td.display(ref / (ref_albedo + eps) * (albedo4 + eps), st.exposure), ref_img)
    log(f"denoised   4 spp: {sec * 1e3:8.1f} ms, PSNR {dn_psnr:.2f} dB")
    dn_ladder = []
    for passes, spps in ((ensemble, denoised_ladder), (2, ensemble_ladder if ensemble != 2 else ())):
        lib.RtSetOption(s.handle, b"denoise_ensemble", passes)
        s.denoise()   # the new setting's buffers before timing
        for spp in spps:
            el, im, _ = denoised(spp)
            dn_ladder.append({"spp": spp, "ensemble": passes, "seconds": round(el, 4),
                              "psnr_db": round(td.psnr(im, ref_img), 3), "tiles_fallback": s.denoise_stats()[1]})
            log(f"denoised {spp:3d} spp x{passes}: {el * 1e3:8.1f} ms, PSNR {dn_ladder[-1]['psnr_db']:.2f} dB")
    lib.RtSetOption(s.handle, b"denoise_ensemble", ensemble)
    hp, wp = (h + 15) // 16 * 16, (w + 15) // 16 * 16
    flops = lib.RtDenoiserForwardFlops(1, hp, wp)
    # equal-PSNR point on the plain ladder (linear in log2 spp, times interpolated alike)
    eq = None
    for a, b in zip(ladder, ladder[1:]):
        if a["psnr_db"] <= dn_psnr <= b["psnr_db"]:
            f = (dn_psnr - a["psnr_db"]) / max(b["psnr_db"] - a["psnr_db"], 1e-9)
            eq = {"spp": round(2 ** (np.log2(a["spp"]) + f * (np.log2(b["spp"]) - np.log2(a["spp"]))), 1),
                  "seconds": a["seconds"] + f * (b["seconds"] - a["seconds"])}
            break
    if eq is None and dn_psnr > ladder[-1]["psnr_db"]:
        eq = {"spp": f">{ladder[-1]['spp']}", "seconds": ladder[-1]["seconds"]}
    p256 = next(x for x in ladder if x["spp"] == 256)
    match = min((d for d in dn_ladder if d["psnr_db"] >= p256["psnr_db"]), key=lambda d: d["seconds"], default=None)
    return {"reference_spp": reference_spp, "reference_seconds": round(ref_s, 2), "denoise_ensemble": ensemble,
            "plain": ladder,
            "denoised_4spp": {"seconds": round(sec, 4), "psnr_db": round(dn_psnr, 3),
                              "plain_4spp_psnr_db": round(td.psnr(plain_img, ref_img), 3),
                              "ceiling_with_4spp_albedo_db": round(ceiling, 3),
                              "tiles": tiles, "tiles_fallback": rejected, "network_ms": round(infer_ms, 3),
                              "network_gflop": round(flops / 1e9, 1),
                              "mfma_utilization": round(flops / (infer_ms / 1e3) / 2.5e15, 4)},
            "equal_psnr_plain": eq,
            "time_to_equal_psnr_speedup": round(eq["seconds"] / sec, 2) if eq else None,
            "vs_256spp": {"psnr_gap_db": round(dn_psnr - p256["psnr_db"], 3),
                          "time_ratio": round(p256["seconds"] / sec, 2)},
            "denoised_ladder": dn_ladder,
            "match_256spp": ({"denoised_spp": match["spp"], "ensemble": match["ensemble"], "seconds": match["seconds"],
                              "time_ratio": round(p256["seconds"] / match["seconds"], 2)} if match else None)}


def psnr_summary(study):
    """The bench line's time_to_equal_psnr: the bar (denoised 4 spp at plain 256-spp PSNR in
    >= 10x less wall time) and the figures behind it, measured in this run."""
    p256 = next(x for x in study["plain"] if x["spp"] == 256)
    d4 = study["denoised_4spp"]
    m = study["match_256spp"]
    return {"measured": "live, in this run, on the bench session's scene",
            "reference_spp": study["reference_spp"],
            "plain_256spp": {"psnr_db": p256["psnr_db"], "seconds": p256["seconds"]},
            "denoised_4spp": {"psnr_db": d4["psnr_db"], "seconds": d4["seconds"], "network_ms": d4["network_ms"],
                              "tiles_fallback": d4["tiles_fallback"], "tiles": d4["tiles"],
                              "ceiling_with_4spp_albedo_db": d4["ceiling_with_4spp_albedo_db"]},
            "denoised_4spp_gap_to_plain_256spp_db": study["vs_256spp"]["psnr_gap_db"],
            "bar_met_at_4spp": bool(study["vs_256spp"]["psnr_gap_db"] >= 0 and study["vs_256spp"]["time_ratio"] >= 10),
            "match_256spp": m,
            "speedup_to_match_256spp": m["time_ratio"] if m else None,
            "speedup_at_equal_psnr_4spp": study["time_to_equal_psnr_speedup"],
            "equal_psnr_plain_spp_4spp": study["equal_psnr_plain"]["spp"] if study["equal_psnr_plain"] else None}


def build_timing(args):
    """--build-timing BUILDS: the GPU BVH build of a workload's scene (its bench launch options),
    BUILDS times in fresh sessions, printed as one JSON object: the workload for profiling the
    build kernels (rocprofv3 --kernel-trace, --pmc) apart from rendering."""
    import native_api
    name = args.workload if args.workload != "batch" else "geodesic"
    desc, tris, cam, st, w, h = workload(name, args.tris)
    cfg = dict(TUNED[name])
    for k in ("bvh_width", "build_max_leaf", "build_bins"):
        if getattr(args, k, None) is not None:
            cfg[k] = getattr(args, k)
    times = []
    for _ in range(args.build_timing):
        with native_api.Session() as s:
            for k, v in cfg.items():
                s.set_option(k, v)
            s.initialize(tris, w, h)
            nodes, depth, ms = s.bvh_info()
            times.append(ms)
        log(f"build {len(times)}: {ms:.1f} ms, {nodes} binary nodes, depth {depth}")
    print(json.dumps({"workload": desc, "triangles": len(tris), "builds": len(times),
                      "build_ms_min": round(min(times), 2), "build_ms_median": round(float(np.median(times)), 2),
                      "binary_nodes": nodes, "depth": depth,
                      "options": {k: cfg[k] for k in ("bvh_width", "build_max_leaf", "build_bins") if k in cfg}}),
          flush=True)
    return 0


def psnr_study(args):
    """--psnr-study: the full study on a C3 workload, printed as one JSON object."""
    import native_api
    name = args.workload if args.workload in C3_WORKLOADS else "geodesic"
    desc, tris, cam, st, w, h = workload(name, 0)
    s = native_api.Session()
    for k, v in TUNED[name].items():
        s.set_option(k, v)   # before initialize: width and leaf size shape the build
    if args.denoise_ensemble:
        s.set_option("denoise_ensemble", args.denoise_ensemble)
    s.initialize(tris, w, h)
    log(f"psnr study: {desc}; reference {args.reference_spp} spp")
    result = {"workload": desc}
    result.update(run_psnr_study(s, cam, st, w, h, args.reference_spp))
    s.close()
    print(json.dumps(result), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="geodesic",
                    choices=("geodesic", "icosphere9", "icosphere", "soup", "cornell", "batch"),
                    help="geodesic: the 7.0M-triangle procedural scene (default); icosphere9 (alias icosphere): "
                         "its 5.24M subdivision-9 icosphere form")
    ap.add_argument("--batch-views", type=int, default=64)
    ap.add_argument("--tris", type=int, default=1_000_000)
    ap.add_argument("--bvh-width", type=int)
    ap.add_argument("--blocks-per-cu", type=int)
    ap.add_argument("--short-stack", type=int)
    ap.add_argument("--refill", type=int)
    ap.add_argument("--refill-below", type=int)
    ap.add_argument("--tail-bounce", type=int,
                    help="bounce from which paths complete in one launch (0: off)")
    ap.add_argument("--pipelines", type=int, choices=(1, 2, 3, 4),
                    help="path pipelines per frame (the library's default is 4)")
    ap.add_argument("--build-max-leaf", type=int)
    ap.add_argument("--build-bins", type=int)
    ap.add_argument("--build-timing", type=int, default=0, metavar="BUILDS",
                    help="only build the workload's BVH BUILDS times (for profiling the build) and exit")
    ap.add_argument("--bvh-quantized", type=int, choices=(0, 1))
    ap.add_argument("--while-while", type=int, choices=(0, 1))
    ap.add_argument("--xcd-segments", type=int, choices=(0, 1),
                    help="per-XCD segment cursors in the traversal launches")
    ap.add_argument("--autotune", action="store_true",
                    help="search the launch configuration now instead of using the recorded tuned one")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--psnr-study", action="store_true",
                    help="only the full time-to-equal-PSNR study of the learned denoiser vs plain accumulation")
    ap.add_argument("--no-psnr", action="store_true",
                    help="skip the live time-to-equal-PSNR study of the C3 workloads (profiling runs)")
    ap.add_argument("--reference-spp", type=int, default=4096)
    ap.add_argument("--denoise-timing", type=int, default=0, metavar="CALLS",
                    help="only time RtDenoise CALLS times on a 4-spp icosphere frame (for profiling) and exit")
    ap.add_argument("--denoise-train-timing", type=int, default=0, metavar="STEPS",
                    help="only time STEPS denoiser training steps on random data (for profiling) and exit")
    ap.add_argument("--train-batch", type=int, default=8, help="crops per training step (--denoise-train-timing)")
    ap.add_argument("--train-crop", type=int, default=256, help="crop size of --denoise-train-timing")
    ap.add_argument("--denoise-ensemble", type=int, choices=(1, 2, 4),
                    help="network passes per denoise (mirrored inputs averaged); default: the library's")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel over BYTES (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (paths are independent per pixel, "
              f"so a multi-GPU run would shard image rows; see the design notes)", file=sys.stderr)
        return 2
    if args.copy_probe:
        import ctypes
        import native_api
        ms = ctypes.c_float()
        if not native_api.load_library().RtBandwidthProbe(args.copy_probe, 5, ctypes.byref(ms)):
            raise SystemExit("RtBandwidthProbe failed")
        gbs = 2 * args.copy_probe / (ms.value / 1e3) / 1e9
        print(json.dumps({"copy_bytes": args.copy_probe, "mean_ms": ms.value, "read_plus_write_GBs": gbs}))
        return 0
    if args.psnr_study:
        return psnr_study(args)
    if args.denoise_timing:
        return denoise_timing(args)
    if args.denoise_train_timing:
        return denoise_train_timing(args)
    if args.build_timing:
        return build_timing(args)
    if args.workload == "batch":
        return bench_batch(args)
    import native_api
    desc, tris, cam, st, w, h = workload(args.workload, args.tris)
    log(f"workload: {desc}")
    cfg = dict(TUNED[args.workload])
    for k in ("bvh_width", "bvh_quantized", "build_max_leaf", "build_bins", "blocks_per_cu", "short_stack", "refill",
              "refill_below", "while_while", "pipelines", "tail_bounce", "xcd_segments"):
        if getattr(args, k) is not None:
            cfg[k] = getattr(args, k)
    if args.autotune:
        import autotune
        t0 = time.perf_counter()
        best, trials = autotune.autotune(tris, cam, st, w, h)
        cfg = {k: v for k, v in best.items() if k != "mrays_per_s"}
        log(f"autotune (library search): {trials} configurations in {time.perf_counter() - t0:.1f} s, best {best}")
    s = native_api.Session()
    for k in ("bvh_width", "bvh_quantized", "build_max_leaf", "build_bins", "blocks_per_cu", "short_stack", "refill",
              "refill_below", "while_while", "pipelines", "tail_bounce", "xcd_segments"):
        if k in cfg:
            s.set_option(k, cfg[k])
    t0 = time.perf_counter()
    s.initialize(tris, w, h)
    init_s = time.perf_counter() - t0
    nodes, depth, build_ms = s.bvh_info()
    log(f"initialize {init_s:.2f} s (BVH {nodes} binary nodes, depth {depth}, build {build_ms:.1f} ms)")

    for _ in range(args.warmup):
        s.render_resident(cam, st)
    log("warmup done")
    rays = 0
    device_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        device_ms += s.render_resident(cam, st)
        c, sh = s.ray_counts()
        rays += c + sh
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = rays / wall / 1e6
    log(f"timed: {args.steps} steps, {ms_per_step:.3f} ms/step wall, {device_ms / args.steps:.3f} ms/step device, "
        f"{value:.1f} Mrays/s")

    # Launch times: a second pass of the same steps and options with every launch bracketed
    # by HIP events on its own stream (stage_timing).  Kept out of the pass above, whose frame
    # rate the event records would lower (by 4% on the icosphere, 23% on the Cornell box).
    s.set_option("stage_timing", 1)
    stage_ms = {}
    busy_ms = 0.0   # union of the traversal launches' intervals (pipelines overlap them)
    for _ in range(args.steps):
        s.render_resident(cam, st)
        for k, (ms, n) in s.stage_times().items():
            a = stage_ms.setdefault(k, [0.0, 0])
            a[0] += ms
            a[1] += n
        busy_ms += s.stage_busy("closest", "shadow", "tail")
    s.set_option("stage_timing", 0)

    # Node and triangle counts: one instrumented frame after the timed region.
    s.set_option("trace_stats", 1)
    s.render_resident(cam, st)
    stats = s.trace_stats()
    closest_rays, shadow_rays = s.ray_counts()
    s.set_option("trace_stats", 0)
    width = s.lib.RtGetOption(s.handle, b"bvh_width")
    quantized = s.lib.RtGetOption(s.handle, b"bvh_quantized")
    node_b = s.lib.RtGetOption(s.handle, b"node_bytes")
    pipelines = s.lib.RtGetOption(s.handle, b"pipelines")   # of the timed frames; the instrumented frame runs one
    # The traversal kernels: each launch traces one bounce's closest-hit rays together with the
    # previous bounce's shadow rays (the "closest" stage), shadow rays alone ("shadow"), or,
    # with tail_bounce, every remaining bounce of the paths ("tail", which also shades them).
    trace_stages = ("closest", "shadow", "tail")
    launches = sum(stage_ms[k][1] for k in trace_stages) / args.steps
    trace_ms = sum(stage_ms[k][0] for k in trace_stages) / args.steps   # summed launch durations
    busy_frame_ms = busy_ms / args.steps
    rays_frame = closest_rays + shadow_rays
    nodes_frame = stats["closest_nodes"] + stats["shadow_nodes"]
    tris_frame = stats["closest_tris"] + stats["shadow_tris"]
    bytes_frame = 48 * rays_frame + nodes_frame * node_b + tris_frame * 48
    bytes_launch = bytes_frame / launches
    mean_launch_s = trace_ms / 1e3 / launches
    # With concurrent pipelines, launches overlap: price the bytes on the time at least one
    # traversal launch was running (per launch: busy time / launches).
    effective_launch_s = busy_frame_ms / 1e3 / launches
    achieved = bytes_launch / effective_launch_s / 1e9
    log(f"stages (ms/frame, summed launch durations): " +
        ", ".join(f"{k} {v[0] / args.steps:.3f} ({v[1] // args.steps} launches)" for k, v in stage_ms.items()))
    for kind, n in (("closest", closest_rays), ("shadow", shadow_rays)):
        log(f"{kind}: {n} rays, {stats[kind + '_nodes'] / max(n, 1):.1f} nodes/ray, "
            f"{stats[kind + '_tris'] / max(n, 1):.1f} tris/ray")
    log(f"traversal: {launches:.0f} launches/frame, {mean_launch_s * 1e3:.3f} ms each, busy {busy_frame_ms:.3f} ms/frame, "
        f"{achieved:.0f} GB/s algorithmic")

    denoise = None
    if args.workload in C3_WORKLOADS and not args.no_psnr:
        log(f"time to equal PSNR (reference {args.reference_spp} spp) ...")
        denoise = psnr_summary(run_psnr_study(s, cam, st, w, h, args.reference_spp, denoised_ladder=(8, 12, 16, 20, 24),
                                              repeats=2))
        log(f"denoised 4 spp {denoise['denoised_4spp']['psnr_db']:.2f} dB vs plain 256 spp "
            f"{denoise['plain_256spp']['psnr_db']:.2f} dB; match of 256 spp: {denoise['match_256spp']}")

    baseline = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        baseline = cpu_baseline(tris, cam, st, w, h, args.cpu_seconds)
        log(f"cpu baseline {baseline['value']:.2f} Mrays/s on {baseline['cores']} threads")

    # Measured HBM traffic of the same launches (rocprof FETCH_SIZE / WRITE_SIZE, calibrated
    # against a plain copy) from the workload's committed profile; rocprof cannot run inside
    # this process.
    traffic = l2_hit = None
    tf = ROOT / "profiles" / f"traffic_{'icosphere9' if args.workload == 'icosphere' else args.workload}.json"
    if tf.exists():
        tj = json.loads(tf.read_text())
        traffic, l2_hit = tj.get("hbm_bytes_per_launch"), tj.get("l2_hit_percent")

    line = {
        "metric": "Mrays/s on 1 MI355X",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": desc, "triangles": int(len(tris)), "width": w, "height": h,
                   "bvh_width": width, "launch": cfg, "autotuned_now": args.autotune,
                   "pipelines": pipelines,
                   "rays_per_step": int(rays / args.steps),
                   "device_ms_per_step": round(device_ms / args.steps, 4), "bvh_build_ms": round(build_ms, 1)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "achieved_basis": "algorithmic bytes per launch (48 B/ray + node visits x node bytes + "
                                       "triangle tests x 48 B) / traversal busy time per launch (union of launch "
                                       "intervals); a rate of algorithmic bytes, most served by L2/MALL, not HBM "
                                       "traffic",
                     "hbm_traffic_GBs": round(traffic / effective_launch_s / 1e9, 1) if traffic else None,
                     "hbm_traffic_frac": round(traffic / effective_launch_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                     "l2_hit_percent": round(l2_hit, 1) if l2_hit is not None else None,
                     "kernel": f"traceKernel<{'QNode' if quantized else 'WideNode'}<{width}>> (closest-hit + shadow queues)"
                               + (" + pathTailKernel" if stage_ms["tail"][1] else ""),
                     "node_bytes": node_b, "mean_launch_ms": round(mean_launch_s * 1e3, 4),
                     "busy_ms_per_step": round(busy_frame_ms, 4),
                     "effective_launch_ms": round(effective_launch_s * 1e3, 4),
                     "launches_per_step": launches, "algorithmic_bytes_per_launch": int(bytes_launch),
                     "nodes_per_ray": round(nodes_frame / max(rays_frame, 1), 2),
                     "tris_per_ray": round(tris_frame / max(rays_frame, 1), 2)},
        "cpu_baseline": baseline,
        "time_to_equal_psnr": denoise,
    }
    s.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
