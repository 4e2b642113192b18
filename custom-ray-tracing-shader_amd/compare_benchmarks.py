#!/usr/bin/env python3
"""Compare two benchmark_native.py runs of the same snapshot and require unchanged images.

Restates Tools/compare_benchmarks.py:7-45: both runs must share the snapshot hash, the
measured and warm-up pass counts, the CPU worker setting and the set of (width, height,
bounces, rays per pixel) cases; per backend, name, sample count and thread count must agree
and the raw RGBA SHA-256 must be identical.  Writes the per-backend speedup table.  A
backend recorded as unavailable in both runs is carried through without timings.

    python compare_benchmarks.py before/ after/ comparison.json
"""
import argparse
import json
from pathlib import Path

METHOD = ("Same scene snapshot, sample counts, bounces and CPU worker count. Per-backend output hashes must match "
          "exactly. Native timings exclude Unity texture upload and display.")


class ComparisonError(AssertionError):
    pass


def _require(condition, message):
    if not condition:
        raise ComparisonError(message)


def case_key(case):
    return (case["width"], case["height"], case["bounces"], case["rays_per_pixel"])


def compare(before, after, log=None):
    """before/after: parsed results.json documents.  Returns the comparison document."""
    _require(before["snapshot_sha256"] == after["snapshot_sha256"], "the two runs rendered different scene snapshots (sha256 mismatch)")
    _require(before["measured_passes"] == after["measured_passes"], "the two runs timed a different number of passes")
    _require(before["warmup_passes"] == after["warmup_passes"], "the two runs used a different number of warm-up passes")
    _require(before["omp_threads"] == after["omp_threads"], "the two runs used different OpenMP thread settings")
    old = {case_key(case): case for case in before["results"]}
    _require(set(old) == {case_key(case) for case in after["results"]}, "the two runs cover different (width, height, bounces, rays per pixel) cases")
    rows = []
    for case in after["results"]:
        previous = old[case_key(case)]
        _require(len(previous["backends"]) == len(case["backends"]), "a case lists a different number of backends in the two runs")
        for a, b in zip(previous["backends"], case["backends"]):
            _require(a["backend"] == b["backend"], f"backend order changed: {a['backend']} vs {b['backend']}")
            if "unavailable" in a or "unavailable" in b:
                _require("unavailable" in a and "unavailable" in b, f"backend {b['backend']} is available in one run only")
                continue
            _require(a["samples_per_pixel"] == b["samples_per_pixel"] and a["cpu_threads"] == b["cpu_threads"],
                     f"backend {b['backend']} rendered a different workload (samples per pixel or CPU threads)")
            _require(a["raw_image_sha256"] == b["raw_image_sha256"],
                     f"output hash of backend {b['backend']} changed for case {case_key(case)}")
            rows.append({"width": case["width"], "height": case["height"], "bounces": case["bounces"],
                         "rays_per_pixel": case["rays_per_pixel"], "samples_per_pixel": b["samples_per_pixel"],
                         "backend": b["backend"], "before_compute_ms": a["compute_median_ms"],
                         "after_compute_ms": b["compute_median_ms"],
                         "compute_speedup": a["compute_median_ms"] / b["compute_median_ms"],
                         "before_call_ms": a["call_median_ms"], "after_call_ms": b["call_median_ms"],
                         "call_speedup": a["call_median_ms"] / b["call_median_ms"], "image_identical": True})
            if log:
                log(f'{b["backend"]} at {case["width"]}x{case["height"]}, {case["bounces"]} bounces: compute median '
                    f'{a["compute_median_ms"]:.3f} ms before, {b["compute_median_ms"]:.3f} ms after '
                    f'(x{a["compute_median_ms"] / b["compute_median_ms"]:.2f}); same output hash')
    return {"snapshot_sha256": before["snapshot_sha256"], "triangles": before["triangles"],
            "measured_passes": before["measured_passes"], "method": METHOD, "results": rows}


def main():
    parser = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument("before", type=Path)
    parser.add_argument("after", type=Path)
    parser.add_argument("output", type=Path)
    args = parser.parse_args()
    before = json.loads((args.before / "results.json").read_text())
    after = json.loads((args.after / "results.json").read_text())
    doc = compare(before, after, log=print)
    args.output.parent.mkdir(parents=True, exist_ok=True)
    args.output.write_text(json.dumps(doc, indent=2) + "\n")


if __name__ == "__main__":
    main()
