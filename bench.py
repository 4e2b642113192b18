"""Mrays/s of the MI355X path tracer on one GPU, with roofline and CPU-baseline figures.

Workload (default): the 1M-triangle random soup at 1920x1080, 3 bounces, 1 progressive
sample per step (no emitters, sky strength 1, so next-event estimation is off), BVH built
once at RtInitialize.  A step is one RtRenderResident call (raygen, traversal, shading and
accumulation of one frame, output left in HBM).  value = (closest-hit + shadow ray queries)
/ wall time of the timed steps.

After the timed region, two instrumented frames measure (a) per-stage kernel time with HIP
events on the session's stream and (b) node visits / triangle tests per ray, from which the
closest-hit traversal kernel's algorithmic bytes per launch are formed:
    bytes = 48 B per ray (32 B ray in + 16 B hit out) + node visits * node bytes
            + triangle tests * 48 B
roofline.achieved = bytes per launch / mean launch time.

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
# Launch configurations chosen by autotune.py on an MI355X for each workload (bench.py
# --autotune repeats the search); the BVH width also fixes the node size below.
# The soup keeps the reference builder's leaf <= 4 rule (its configuration names it); the
# icosphere uses the autotuner's leaf size.
TUNED = {
    "soup": {"bvh_width": 4, "build_max_leaf": 2, "blocks_per_cu": 2, "short_stack": 12, "while_while": 1, "refill": 1,
             "pipelines": 4},
    "icosphere": {"bvh_width": 4, "build_max_leaf": 2, "blocks_per_cu": 3, "short_stack": 12, "while_while": 1,
                  "refill": 1, "pipelines": 4, "tail_bounce": 4},
    "cornell": {"bvh_width": 4, "build_max_leaf": 4, "blocks_per_cu": 8, "short_stack": 12, "while_while": 1, "refill": 0,
                "pipelines": 4},
    "batch": {"bvh_width": 4, "blocks_per_cu": 3, "short_stack": 12, "while_while": 1, "refill": 1, "pipelines": 3},
}


def batch_views(count, seed=2024):
    """Batched views: scenes from the soup / displaced-sphere generators with triangle counts
    log-uniform in [1e3, 2e6], image sides log-uniform in [256, 1024], 2..8 bounces."""
    import scenes
    rng = np.random.default_rng(seed)
    out = []
    for v in range(count):
        n = int(round(10 ** rng.uniform(3, np.log10(2e6))))
        w, h = (int(round(2 ** rng.uniform(8, 10))) for _ in range(2))
        bounces = int(rng.integers(2, 9))
        if v % 2 == 0:
            tris = scenes.triangle_soup(n, seed=seed + v)
            cam, st = scenes.soup_camera(w, h), scenes.make_settings(bounces, sky=1.0, seed=seed + v)
        else:
            sub = int(np.clip(round(np.log(max(n - 20, 20) / 20) / np.log(4)), 1, 8))
            tris = scenes.displaced_icosphere(subdiv=sub, seed=seed + v)
            cam, st = scenes.icosphere_camera(w, h), scenes.make_settings(bounces, seed=seed + v)
        out.append((tris, w, h, cam, st))
    return out


def bench_batch(args):
    import native_api
    views = batch_views(args.batch_views)
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
        log(f"autotune: {len(trials)} batch configurations in {time.perf_counter() - t0:.1f} s, best {best}")
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
        return (f"random triangle soup {len(tris)} tris, 1920x1080, 3 bounces, 1 spp/step, sky 1",
                tris, scenes.soup_camera(w, h), scenes.make_settings(3, rpp=1, sky=1.0), w, h)
    if name == "icosphere":
        w, h = 1920, 1080
        tris = scenes.displaced_icosphere(subdiv=9, seed=7)
        return (f"displaced icosphere + room {len(tris)} tris, 1920x1080, 8 bounces + RR, 1 spp/step",
                tris, scenes.icosphere_camera(w, h), scenes.make_settings(8, rpp=1, sky=0.0), w, h)
    if name == "cornell":
        w = h = 256
        return ("Cornell box 36 tris, 256x256, 4 bounces, 16 spp/step", scenes.cornell(), scenes.cornell_camera(w, h),
                scenes.make_settings(4, rpp=16), w, h)
    raise SystemExit(f"unknown workload {name}")


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
            "sample": f"oracle (C++ restatement of the reference CPU path, 12-bin SAH BVH, OpenMP over rows, "
                      f"{threads} threads) on {rows} rows x {w} px of the same frame and settings: "
                      f"{rays} ray queries in {elapsed:.1f} s"}


def denoise_timing(args):
    """RtDenoise on a 4-spp 1080p frame of the icosphere workload, repeated: the workload for
    profiling the denoiser's kernels (rocprofv3 --kernel-trace) apart from the study."""
    import native_api
    desc, tris, cam, st, w, h = workload("icosphere", 0)
    s = native_api.Session()
    for k, v in TUNED["icosphere"].items():
        s.set_option(k, v)
    s.initialize(tris, w, h)
    st.raysPerPixel = 4
    s.render(cam, st)
    times = []
    for _ in range(args.denoise_timing):
        _, ms = s.denoise()
        times.append(ms)
    tiles, rejected, infer_ms = s.denoise_stats()
    flops = s.lib.RtDenoiserForwardFlops(1, (h + 15) // 16 * 16, (w + 15) // 16 * 16)
    s.close()
    print(json.dumps({"workload": desc + ", 4 spp", "denoise_calls": len(times),
                      "rtdenoise_ms_min": round(min(times), 3), "network_ms_last": round(infer_ms, 3),
                      "network_flop": flops, "tiles": tiles, "tiles_fallback": rejected}), flush=True)
    return 0


def psnr_study(args):
    """Time to equal PSNR: learned denoise at 4 spp vs plain accumulation, against a
    high-sample render (seed 20261023), 8-bit display PSNR as Tools/verify_quality.py:53-54.
    Wall times include each RtRender's readback (Tools/verify_quality.py:72-78) and the
    denoiser's transfers."""
    import ctypes
    import native_api
    import train_denoiser as td
    desc, tris, cam, st, w, h = workload("icosphere", 0)
    lib = native_api.load_library()
    s = native_api.Session(lib)
    for k, v in TUNED["icosphere"].items():
        s.set_option(k, v)   # before initialize: width and leaf size shape the build
    if args.denoise_ensemble:
        s.set_option("denoise_ensemble", args.denoise_ensemble)
    ensemble = s.lib.RtGetOption(s.handle, b"denoise_ensemble")
    s.initialize(tris, w, h)
    log(f"psnr study: {desc}; reference {args.reference_spp} spp")
    t0 = time.perf_counter()
    ref = td.render_mean(s, cam, st, args.reference_spp, seed=20261023)
    ref_s = time.perf_counter() - t0
    ref_img = td.display(ref, st.exposure)

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
        img, sec = plain(spp)
        ladder.append({"spp": spp, "seconds": round(sec, 4), "psnr_db": round(td.psnr(img, ref_img), 3)})
        log(f"plain {spp:4d} spp: {sec * 1e3:8.1f} ms, PSNR {ladder[-1]['psnr_db']:.2f} dB")
    plain(4)          # warm the denoiser path once (weights load, buffers)
    s.denoise()
    denoised = []
    for _ in range(3):
        t = time.perf_counter()
        plain_img, _ = plain(4)
        img, _ = s.denoise()
        denoised.append((time.perf_counter() - t, img))
    sec, img = min(denoised, key=lambda x: x[0])
    tiles, rejected, infer_ms = s.denoise_stats()
    dn_psnr = td.psnr(img, ref_img)
    # Denoised frames at more samples: the smallest count whose denoised PSNR reaches plain
    # accumulation at 256 spp, and how much sooner it gets there.
    dn_ladder = []
    for spp in (8, 12, 16, 20, 24, 32, 64):
        best = None
        for _ in range(2):
            t = time.perf_counter()
            plain(spp)
            im, _ = s.denoise()
            el = time.perf_counter() - t
            best = el if best is None else min(best, el)
        dn_ladder.append({"spp": spp, "seconds": round(best, 4), "psnr_db": round(td.psnr(im, ref_img), 3),
                          "tiles_fallback": s.denoise_stats()[1]})
        log(f"denoised {spp:3d} spp: {best * 1e3:8.1f} ms, PSNR {dn_ladder[-1]['psnr_db']:.2f} dB")
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
    result = {"workload": desc, "reference_spp": args.reference_spp, "reference_seconds": round(ref_s, 2),
              "denoise_ensemble": ensemble,
              "plain": ladder, "denoised_4spp": {"seconds": round(sec, 4), "psnr_db": round(dn_psnr, 3),
                                                 "plain_4spp_psnr_db": round(td.psnr(plain_img, ref_img), 3),
                                                 "tiles": tiles, "tiles_fallback": rejected,
                                                 "network_ms": round(infer_ms, 3),
                                                 "network_gflop": round(flops / 1e9, 1),
                                                 "mfma_utilization": round(flops / (infer_ms / 1e3) / 2.5e15, 4)},
              "equal_psnr_plain": eq,
              "time_to_equal_psnr_speedup": round(eq["seconds"] / sec, 2) if eq else None,
              "vs_256spp": {"psnr_gap_db": round(dn_psnr - p256["psnr_db"], 3),
                            "time_ratio": round(p256["seconds"] / sec, 2)},
              "denoised_ladder": dn_ladder}
    match = next((d for d in dn_ladder if d["psnr_db"] >= p256["psnr_db"]), None)
    result["match_256spp"] = ({"denoised_spp": match["spp"], "seconds": match["seconds"],
                               "time_ratio": round(p256["seconds"] / match["seconds"], 2)} if match else None)
    s.close()
    print(json.dumps(result), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="soup", choices=("soup", "icosphere", "cornell", "batch"))
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
                    help="path pipelines per frame (the library's default is 2)")
    ap.add_argument("--build-max-leaf", type=int)
    ap.add_argument("--bvh-quantized", type=int, choices=(0, 1))
    ap.add_argument("--while-while", type=int, choices=(0, 1))
    ap.add_argument("--xcd-segments", type=int, choices=(0, 1),
                    help="per-XCD segment cursors in the traversal launches")
    ap.add_argument("--autotune", action="store_true",
                    help="search the launch configuration now instead of using the recorded tuned one")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--psnr-study", action="store_true",
                    help="time-to-equal-PSNR of the learned denoiser vs plain accumulation on the icosphere")
    ap.add_argument("--reference-spp", type=int, default=4096)
    ap.add_argument("--denoise-timing", type=int, default=0, metavar="CALLS",
                    help="only time RtDenoise CALLS times on a 4-spp icosphere frame (for profiling) and exit")
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
    if args.workload == "batch":
        return bench_batch(args)
    import native_api
    desc, tris, cam, st, w, h = workload(args.workload, args.tris)
    log(f"workload: {desc}")
    cfg = dict(TUNED[args.workload])
    for k in ("bvh_width", "bvh_quantized", "build_max_leaf", "blocks_per_cu", "short_stack", "refill", "refill_below",
              "while_while", "pipelines", "tail_bounce", "xcd_segments"):
        if getattr(args, k) is not None:
            cfg[k] = getattr(args, k)
    if args.autotune:
        import autotune
        t0 = time.perf_counter()
        best, trials = autotune.autotune(tris, cam, st, w, h, use_cache=False,
                                         max_leaf=(2, 4) if args.workload == "soup" else (2, 3, 4))
        cfg = {k: v for k, v in best.items() if k != "mrays_per_s"}
        log(f"autotune: {len(trials)} configurations in {time.perf_counter() - t0:.1f} s, best {best}")
    s = native_api.Session()
    for k in ("bvh_width", "bvh_quantized", "build_max_leaf", "blocks_per_cu", "short_stack", "refill", "refill_below",
              "while_while", "pipelines", "tail_bounce", "xcd_segments"):
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

    baseline = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        baseline = cpu_baseline(tris, cam, st, w, h, args.cpu_seconds)
        log(f"cpu baseline {baseline['value']:.2f} Mrays/s on {baseline['cores']} threads")

    traffic = None
    tf = ROOT / "profiles" / f"traffic_{args.workload}.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get("hbm_bytes_per_launch")
    denoise = None
    pf = ROOT / "profiles" / "denoise_psnr_icosphere.json"
    if pf.exists():
        d = json.loads(pf.read_text())
        denoise = {"source": "profiles/denoise_psnr_icosphere.json (bench.py --psnr-study)",
                   "time_to_equal_psnr_speedup": d.get("time_to_equal_psnr_speedup"),
                   "denoised_4spp_psnr_db": d["denoised_4spp"]["psnr_db"],
                   "plain_256spp_psnr_db": next(x["psnr_db"] for x in d["plain"] if x["spp"] == 256),
                   "equal_psnr_plain_spp": d["equal_psnr_plain"]["spp"] if d.get("equal_psnr_plain") else None}

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
