"""Per-layer roofline of one denoiser inference from a rocprofv3 --kernel-trace run of
`bench.py --denoise-timing N`: each layer's measured time beside its attainable time, the
larger of its MFMA FLOP at the dense BF16 peak (2.5 PFLOP/s) and its HBM bytes at 8 TB/s.

Bytes count each tensor once (halo re-reads are L2 hits): the layer's input channels at its
resolution (the upsampled channels of a decoder's first convolution are read in place at
half resolution, a quarter of the pixels), its output channels, BF16; the last layer reads
BF16 and writes 4 FP32 channels.  FLOP are 2 x 9 x input channels x output channels per
pixel, with the channel counts the kernels run (multiples of 32; 16 output channels for the
3-channel last layer).  The launches of the last call are matched in order: the feature
kernel, then the network's layers and pooling, then the output kernels.
"""
import argparse
import csv
import glob
import os

PEAK_FLOPS, PEAK_BYTES = 2.5e15, 8.0e12
# (name, level, input channels run, output channels run, upsampled input channels read at
# half resolution, output FP32 channels instead of BF16)
LAYERS = [("E0A", 0, 32, 32, 0, 0), ("E0B", 0, 32, 32, 0, 0), ("pool0", 0, 32, 32, 0, 0),
          ("E1A", 1, 32, 64, 0, 0), ("E1B", 1, 64, 64, 0, 0), ("pool1", 1, 64, 64, 0, 0),
          ("E2A", 2, 64, 64, 0, 0), ("E2B", 2, 64, 64, 0, 0), ("pool2", 2, 64, 64, 0, 0),
          ("E3A", 3, 64, 96, 0, 0), ("E3B", 3, 96, 96, 0, 0), ("pool3", 3, 96, 96, 0, 0),
          ("BA", 4, 96, 128, 0, 0), ("BB", 4, 128, 128, 0, 0),
          ("D3A", 3, 224, 96, 128, 0), ("D3B", 3, 96, 96, 0, 0), ("D2A", 2, 160, 64, 96, 0),
          ("D2B", 2, 64, 64, 0, 0), ("D1A", 1, 128, 64, 64, 0), ("D1B", 1, 64, 64, 0, 0),
          ("D0A", 0, 96, 32, 64, 0), ("D0B", 0, 32, 32, 0, 0), ("OUT", 0, 32, 16, 0, 4)]


def layer_cost(layer, width, height):
    name, level, cin, cout, up, f32 = layer
    px = (width >> level) * (height >> level)
    if name.startswith("pool"):
        return 0.0, px * cin * 2 + px // 4 * cin * 2
    flop = 2.0 * 9 * cin * cout * px
    read = (cin - up) * 2 * px + up * 2 * px / 4
    write = f32 * 4 * px if f32 else cout * 2 * px
    return flop, read + write


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--trace", required=True, help="rocprofv3 output directory (csv)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1088)
    ap.add_argument("--command", required=True, help="the command that made the trace")
    ap.add_argument("--md", required=True)
    args = ap.parse_args()
    files = glob.glob(os.path.join(args.trace, "**", "*kernel_trace.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no kernel_trace.csv under {args.trace}")
    rows = sorted(csv.DictReader(open(files[0])), key=lambda r: int(r["Start_Timestamp"]))
    starts = [i for i, r in enumerate(rows) if "featuresKernel" in r["Kernel_Name"]]
    if not starts:
        raise SystemExit("no featuresKernel launch in the trace")
    seq = [r for r in rows[starts[-1] + 1:] if "conv" in r["Kernel_Name"] or "Pool" in r["Kernel_Name"]]
    if len(seq) < len(LAYERS):
        raise SystemExit(f"{len(seq)} network launches after the last feature kernel, expected {len(LAYERS)}")
    out = [f"# denoiser inference at {args.width}x{args.height}: per-layer roofline", "",
           "Command (from the repository root, on one MI355X):", "", "```", args.command, "```", "",
           "| layer | kernel | us | GFLOP | HBM MB | attainable us | bound | fraction |",
           "|---|---|---:|---:|---:|---:|---|---:|"]
    tot_t = tot_a = tot_f = 0.0
    for layer, r in zip(LAYERS, seq):
        us = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        flop, byts = layer_cost(layer, args.width, args.height)
        t_flop, t_byte = flop / PEAK_FLOPS * 1e6, byts / PEAK_BYTES * 1e6
        att = max(t_flop, t_byte)
        kernel = r["Kernel_Name"].replace("rt::", "").replace("(anonymous namespace)::", "").split("(")[0]
        out.append(f"| {layer[0]} | `{kernel}` | {us:.1f} | {flop / 1e9:.1f} | {byts / 1e6:.0f} | {att:.1f} | "
                   f"{'MFMA' if t_flop >= t_byte else 'HBM'} | {att / us:.2f} |")
        tot_t += us; tot_a += att; tot_f += flop
    out += ["", f"Network: {tot_t:.0f} us of launch durations against {tot_a:.0f} us attainable "
                f"({tot_a / tot_t:.2f}); {tot_f / 1e9:.0f} GFLOP at {tot_f / (tot_t * 1e-6) / 1e12:.0f} TFLOP/s, "
                f"{tot_f / (tot_t * 1e-6) / PEAK_FLOPS:.3f} of the dense BF16 peak."]
    open(args.md, "w").write("\n".join(out) + "\n")
    print("\n".join(out))


if __name__ == "__main__":
    main()
