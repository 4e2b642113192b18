"""Per-kernel summary of a rocprofv3 --kernel-trace --stats run, for kernels matching a
pattern, per call of the profiled workload (e.g. per RtDenoise or per BVH build), with the
MFMA utilisation of a given FLOP count over the matched kernels' time (dense BF16 peak
2.5 PFLOP/s) and, from two rocprofv3 --pmc runs of the same workload (FETCH_SIZE, then
WRITE_SIZE, in KiB per dispatch), each kernel's HBM bytes per launch, scaled by what the
counters read for a plain copy kernel of known size."""
import argparse
import csv
import glob
import os
import re


def counter_means(directory):
    """Mean counter bytes per dispatch (the counters count KiB) by kernel name."""
    if not directory:
        return {}
    files = glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {directory}")
    sums, counts = {}, {}
    for r in csv.DictReader(open(files[0])):
        k = r["Kernel_Name"]
        sums[k] = sums.get(k, 0.0) + float(r["Counter_Value"]) * 1024
        counts[k] = counts.get(k, 0) + 1
    return {k: sums[k] / counts[k] for k in sums}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--trace", required=True, help="rocprofv3 output directory (csv)")
    ap.add_argument("--match", required=True, help="regex over kernel names")
    ap.add_argument("--calls", type=int, required=True, help="workload calls in the profiled run")
    ap.add_argument("--flop", type=float, default=0.0, help="FLOP of one workload call on MFMA")
    ap.add_argument("--mfma-match", default="conv", help="regex of the kernels the FLOP run on")
    ap.add_argument("--title", required=True)
    ap.add_argument("--command", required=True, help="the command that made the trace")
    ap.add_argument("--md", required=True)
    ap.add_argument("--fetch", help="rocprofv3 --pmc FETCH_SIZE output directory (csv) of the workload")
    ap.add_argument("--write", help="rocprofv3 --pmc WRITE_SIZE output directory (csv) of the workload")
    ap.add_argument("--fetch-scale", type=float, default=1.0,
                    help="HBM bytes read per FETCH_SIZE byte (copy kernel calibration)")
    ap.add_argument("--write-scale", type=float, default=1.0,
                    help="HBM bytes written per WRITE_SIZE byte (copy kernel calibration)")
    ap.add_argument("--counter-commands", default="", help="the commands that made the counter runs")
    args = ap.parse_args()
    stats = glob.glob(os.path.join(args.trace, "**", "*kernel_stats.csv"), recursive=True)
    if not stats:
        raise SystemExit(f"no kernel_stats.csv under {args.trace}")
    pat, mpat = re.compile(args.match), re.compile(args.mfma_match)
    rows = [r for r in csv.DictReader(open(stats[0])) if pat.search(r["Name"])]
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    total = sum(float(r["TotalDurationNs"]) for r in rows) / args.calls / 1e3
    mfma_us = sum(float(r["TotalDurationNs"]) for r in rows if mpat.search(r["Name"])) / args.calls / 1e3
    fetch, write = counter_means(args.fetch), counter_means(args.write)
    traffic = bool(fetch or write)
    out = [f"# {args.title}", "", "Command (from the repository root, on one MI355X):", "", "```", args.command, "```", ""]
    if traffic:
        out += ["HBM bytes per launch from the counter runs (FETCH_SIZE x %.4g, WRITE_SIZE x %.4g, the factors that "
                "make a plain copy kernel's counters equal its bytes):" % (args.fetch_scale, args.write_scale), "",
                "```", args.counter_commands, "```", "",
                "| kernel | launches / call | mean us | us / call | HBM read MB / launch | HBM write MB / launch |",
                "|---|---:|---:|---:|---:|---:|"]
    else:
        out += ["| kernel | launches / call | mean us | us / call |", "|---|---:|---:|---:|"]
    for r in rows:
        name = re.sub(r"rt::|\(anonymous namespace\)::", "", r["Name"])
        name = name if len(name) <= 72 else name[:69] + "..."
        per = float(r["TotalDurationNs"]) / args.calls / 1e3
        line = f"| `{name}` | {int(r['Calls']) / args.calls:.1f} | {float(r['AverageNs']) / 1e3:.1f} | {per:.1f} |"
        if traffic:
            f, w = fetch.get(r["Name"]), write.get(r["Name"])
            line += (f" {f * args.fetch_scale / 1e6:.1f} |" if f is not None else " - |")
            line += (f" {w * args.write_scale / 1e6:.1f} |" if w is not None else " - |")
        out.append(line)
    out += ["", f"Matched kernels: {total:.1f} us of summed launch durations per call (launches on concurrent "
                "streams overlap, so this can exceed the call's wall time)."]
    if args.flop and mfma_us:
        rate = args.flop / (mfma_us * 1e-6)
        out.append(f"Kernels on MFMA ({args.mfma_match}): {mfma_us:.1f} us per call for {args.flop / 1e9:.1f} GFLOP: "
                   f"{rate / 1e12:.0f} TFLOP/s, {rate / 2.5e15:.3f} of the 2.5 PFLOP/s dense BF16 peak.")
    open(args.md, "w").write("\n".join(out) + "\n")
    print("\n".join(out))


if __name__ == "__main__":
    main()
