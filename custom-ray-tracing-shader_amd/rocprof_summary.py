"""Summarise rocprofv3 output of a bench.py run into a short Markdown table plus a JSON
with the traversal kernel's HBM bytes per launch (read by bench.py as roofline.traffic).

Inputs (directories written by rocprofv3 -d ... --output-format csv):
  --trace DIR        --kernel-trace --stats run            (*_kernel_stats.csv, *_kernel_trace.csv)
  --fetch DIR        --pmc FETCH_SIZE run of the workload   (*_counter_collection.csv)
  --write DIR        --pmc WRITE_SIZE run of the workload
  --tcc DIR          --pmc TCC_HIT_sum TCC_MISS_sum run     (optional)
  --sq DIR           --pmc SQ_WAVES SQ_INSTS_VALU SQ_INSTS_VMEM_RD SQ_INSTS_LDS SQ_ACTIVE_INST_VALU
                     SQ_THREAD_CYCLES_VALU run (optional): instruction mix and VALU lane use
  --copy-fetch DIR / --copy-write DIR   the same counters over bench.py --copy-probe BYTES
  --copy-bytes BYTES                    size the probe copied (reads BYTES, writes BYTES)
FETCH_SIZE and WRITE_SIZE are in KiB per dispatch; each is multiplied by
(expected KiB of the copy) / (counter KiB of the copy) before use.

Launches of concurrent path pipelines overlap, so the dominant kernel's time per launch is
also priced as the union of its launch intervals (Start_Timestamp..End_Timestamp of every
matching dispatch in the kernel trace) divided by the launches: with --bench-json, the
algorithmic bytes per launch over that time give the roofline fraction from this trace alone.
"""
import argparse
import csv
import glob
import json
import os
import re
import sys
from collections import defaultdict


def one_csv(directory, suffix):
    files = sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True))
    if not files:
        raise SystemExit(f"no *{suffix} under {directory}")
    return list(csv.DictReader(open(files[0])))


def busy_union_ns(rows, pattern):
    """(launches, union of their [start, end) intervals in ns, summed durations in ns)."""
    spans = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in rows
                   if pattern.search(r["Kernel_Name"]))
    busy, end, total = 0, None, 0
    for a, b in spans:
        total += b - a
        if end is None or a > end:
            busy += b - a
            end = b
        elif b > end:
            busy += b - end
            end = b
    return len(spans), busy, total


def counter_means(directory):
    """{kernel name: {counter: mean value per dispatch}}"""
    acc = defaultdict(lambda: defaultdict(list))
    for r in one_csv(directory, "counter_collection.csv"):
        acc[r["Kernel_Name"]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in d.items()} for k, d in acc.items()}


def short(name, n=72):
    name = name.replace("rt::(anonymous namespace)::", "").replace("rt::", "")
    return name if len(name) <= n else name[:n - 3] + "..."


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--tcc")
    ap.add_argument("--sq")
    ap.add_argument("--copy-fetch", required=True)
    ap.add_argument("--copy-write", required=True)
    ap.add_argument("--copy-bytes", type=int, required=True)
    ap.add_argument("--kernel", required=True,
                    help="regular expression naming the dominant kernel; launches of every matching "
                         "instantiation are pooled (means weighted by calls)")
    ap.add_argument("--title", required=True)
    ap.add_argument("--commands", required=True, help="file holding the commands that made the inputs")
    ap.add_argument("--bench-json", help="bench.py output line of the same workload (algorithmic bytes)")
    ap.add_argument("--md", required=True)
    ap.add_argument("--json", required=True)
    args = ap.parse_args()

    stats = one_csv(args.trace, "kernel_stats.csv")
    fetch, write = counter_means(args.fetch), counter_means(args.write)
    cf, cw = counter_means(args.copy_fetch), counter_means(args.copy_write)
    copy_name = next(k for k in cf if "copyKernel" in k)
    expected_kib = args.copy_bytes / 1024
    fetch_scale = expected_kib / cf[copy_name]["FETCH_SIZE"]
    write_scale = expected_kib / cw[copy_name]["WRITE_SIZE"]
    tcc = counter_means(args.tcc) if args.tcc else {}

    lines = [f"# {args.title}", "", "Commands (run from the repository root on one MI355X):", "", "```",
             open(args.commands).read().strip(), "```", "",
             f"HBM counter calibration with a plain {args.copy_bytes / 2**30:.2f} GiB device copy: "
             f"FETCH_SIZE x {fetch_scale:.3f}, WRITE_SIZE x {write_scale:.3f}.", "",
             "| kernel | calls | mean us | % time | HBM read MB/launch | HBM write MB/launch | L2 hit % |",
             "|---|---:|---:|---:|---:|---:|---:|"]
    pat = re.compile(args.kernel)
    pool = []   # (name, calls, mean ns, read B, write B, hit %) of the matching kernels
    for r in stats:
        name = r["Name"]
        if pat.search(name):
            f = fetch.get(name, {}).get("FETCH_SIZE")
            w = write.get(name, {}).get("WRITE_SIZE")
            t = tcc.get(name, {})
            pool.append((name, int(r["Calls"]), float(r["AverageNs"]),
                         f * fetch_scale * 1024 if f is not None else None,
                         w * write_scale * 1024 if w is not None else None,
                         (t["TCC_HIT_sum"], t["TCC_HIT_sum"] + t["TCC_MISS_sum"]) if t else None))
    for r in sorted(stats, key=lambda r: -float(r["Percentage"]))[:14]:
        name = r["Name"]
        f = fetch.get(name, {}).get("FETCH_SIZE")
        w = write.get(name, {}).get("WRITE_SIZE")
        t = tcc.get(name, {})
        hit = 100 * t["TCC_HIT_sum"] / max(t["TCC_HIT_sum"] + t["TCC_MISS_sum"], 1) if t else None
        rd = f * fetch_scale * 1024 / 1e6 if f is not None else None
        wr = w * write_scale * 1024 / 1e6 if w is not None else None
        fmt = lambda v, p=1: "—" if v is None else f"{v:.{p}f}"
        lines.append(f"| `{short(name)}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} "
                     f"| {fmt(rd)} | {fmt(wr)} | {fmt(hit)} |")
    if not pool:
        raise SystemExit(f"kernel {args.kernel!r} not in the trace")
    calls = sum(p[1] for p in pool)
    wmean = lambda i: (sum(p[1] * p[i] for p in pool) / calls) if all(p[i] is not None for p in pool) else None
    hits = [p[5] for p in pool if p[5]]
    launches, busy_ns, summed_ns = busy_union_ns(one_csv(args.trace, "kernel_trace.csv"), pat)
    dominant = {"kernel": " + ".join(p[0] for p in pool), "kernel_pattern": args.kernel,
                "trace_launches": launches, "trace_busy_ms": busy_ns / 1e6,
                "trace_union_ms_per_launch": busy_ns / 1e6 / max(launches, 1),
                "mean_ms": wmean(2) / 1e6, "calls": calls,
                "hbm_read_bytes_per_launch": wmean(3), "hbm_write_bytes_per_launch": wmean(4),
                "l2_hit_percent": 100 * sum(h[0] for h in hits) / max(sum(h[1] for h in hits), 1) if hits else None}
    traffic = (dominant["hbm_read_bytes_per_launch"] or 0) + (dominant["hbm_write_bytes_per_launch"] or 0)
    dominant["hbm_bytes_per_launch"] = traffic
    dominant["fetch_scale"], dominant["write_scale"] = fetch_scale, write_scale
    lines += ["", f"Kernel trace: {launches} matching launches, summed durations {summed_ns / 1e6:.3f} ms, union of "
              f"their intervals {busy_ns / 1e6:.3f} ms ({busy_ns / 1e3 / max(launches, 1):.1f} us per launch; "
              f"{summed_ns / max(busy_ns, 1):.2f} launches running on average while any runs)."]
    lines += ["", f"Dominant kernel (all launches matching `{args.kernel}`: "
              + ", ".join(f"`{short(p[0], 90)}` x{p[1]}" for p in pool)
              + f"): mean {dominant['mean_ms'] * 1e3:.1f} us over {calls} launches; HBM traffic "
              f"{traffic / 1e6:.1f} MB per launch ({traffic / dominant['mean_ms'] / 1e6:.0f} GB/s of HBM)."]
    if args.sq:
        sqm = counter_means(args.sq)
        sqs = [(p[1], sqm[p[0]]) for p in pool if p[0] in sqm]
        sq = ({c: sum(n * d[c] for n, d in sqs) / sum(n for n, _ in sqs) for c in sqs[0][1]} if sqs else None)
        if sq:
            waves = max(sq["SQ_WAVES"], 1)
            util = sq["SQ_THREAD_CYCLES_VALU"] / max(sq["SQ_ACTIVE_INST_VALU"] * 64, 1)
            dominant.update({"valu_insts_per_wave": sq["SQ_INSTS_VALU"] / waves,
                             "vmem_reads_per_wave": sq["SQ_INSTS_VMEM_RD"] / waves,
                             "lds_insts_per_wave": sq["SQ_INSTS_LDS"] / waves,
                             "valu_lane_utilization": util})
            lines += [f"Instruction mix per wave: {sq['SQ_INSTS_VALU'] / waves:.0f} VALU, "
                      f"{sq['SQ_INSTS_VMEM_RD'] / waves:.0f} vector-memory reads, {sq['SQ_INSTS_LDS'] / waves:.0f} LDS; "
                      f"VALU lane utilization (SQ_THREAD_CYCLES_VALU / (64 x SQ_ACTIVE_INST_VALU)) {100 * util:.1f}%."]
    if args.bench_json:
        b = json.loads(open(args.bench_json).read().strip().splitlines()[-1])
        rf = b["roofline"]
        alg = rf["algorithmic_bytes_per_launch"]
        lines += [f"bench.py measured {rf['mean_launch_ms'] * 1e3:.1f} us per launch with HIP events; algorithmic "
                  f"bytes {alg / 1e6:.1f} MB per launch ({rf['nodes_per_ray']} nodes and {rf['tris_per_ray']} "
                  f"triangles per ray), i.e. {alg / max(traffic, 1):.1f}x the HBM traffic: the rest is served "
                  f"by L2 / MALL."]
        dominant["algorithmic_bytes_per_launch"] = alg
        per_launch_s = busy_ns / 1e9 / max(launches, 1)
        achieved = alg / per_launch_s / 1e9
        dominant["trace_achieved_GBs"] = achieved
        dominant["trace_frac"] = achieved / 8000.0
        dominant["trace_hbm_traffic_frac"] = traffic / per_launch_s / 1e9 / 8000.0
        lines += [f"Roofline from this trace: algorithmic bytes per launch / union time per launch = "
                  f"{alg / 1e6:.1f} MB / {per_launch_s * 1e6:.1f} us = {achieved:.0f} GB/s, "
                  f"**{achieved / 8000:.3f} of 8 TB/s** (a rate of algorithmic bytes); measured HBM traffic over "
                  f"the same time {traffic / per_launch_s / 1e9:.0f} GB/s, {traffic / per_launch_s / 1e9 / 8000:.3f} "
                  f"of 8 TB/s."]
        if "busy_ms_per_step" in rf:
            lines += [f"bench.py's own figure (HIP events, union per step {rf['busy_ms_per_step']:.3f} ms over "
                      f"{rf['launches_per_step']:.0f} launches, {rf['effective_launch_ms'] * 1e3:.1f} us per launch): "
                      f"{rf['achieved']:.0f} GB/s, {rf['frac']:.3f}."]
    lines += ["", "Summary made with:", "```", "python3 custom-ray-tracing-shader_amd/rocprof_summary.py " +
              " ".join(f'"{a}"' if (" " in a or "|" in a) else a for a in sys.argv[1:]), "```"]
    with open(args.md, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    with open(args.json, "w") as fh:
        json.dump(dominant, fh, indent=1)
    print("\n".join(lines))


if __name__ == "__main__":
    main()
