"""Summaries of rocprofv3 CSV output as the short text tables kept under profiles/.

  python -m hexed_amd.rocprof_summary stats <kernel_stats.csv> --title T --command C > out.md
  python -m hexed_amd.rocprof_summary traffic <FETCH_SIZE counter_collection.csv> <WRITE_SIZE counter_collection.csv>
         --key K --title T --command C [--json profiles/traffic.json] > out.md
  python -m hexed_amd.rocprof_summary counters <counter_collection.csv>... --stats <kernel_stats.csv> --title T
         --command C > out.md

`counters` averages every counter of the CSVs per dispatch and kernel; with SQ_INSTS_MFMA it also gives each kernel's
matrix-core rate: V_MFMA_F32_4X4X1_16B_F32 (the only matrix instruction of these kernels) does 16 4x4x1 products per
wavefront, 512 FLOP, and the rate is that over the kernel's average duration from the kernel-trace statistics, against
the 157.3 TFLOP/s FP32 matrix peak of an MI355X.

`traffic` turns FETCH_SIZE / WRITE_SIZE (KiB per dispatch) into HBM bytes per launch. It scales each counter by what
it reads for hx_stream_copy, the library's plain copy kernel of known size. The copy kernel must have run in the same
profiled process (bench.py runs it): it moves 2 GiB each way per dispatch.
"""
import argparse
import collections
import csv
import json
import os
import sys

COPY_KERNEL = "k_stream_copy"
COPY_BYTES = 2 ** 31
MFMA_FLOP = 512          # V_MFMA_F32_4X4X1_16B_F32 per wavefront
FP32_MATRIX_PEAK = 157.3e12


def _kernel(name):
    name = name.replace("(anonymous namespace)::", "")
    return name.split("(")[0].replace("void ", "").replace("hx::", "") if name.startswith("void ") else name


def stats_table(path):
    rows = list(csv.DictReader(open(path)))
    out = []
    for r in rows:
        out.append((_kernel(r["Name"]), int(r["Calls"]), float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3,
                    float(r["MaxNs"]) / 1e3, float(r["Percentage"])))
    out.sort(key=lambda x: -x[5])
    return out


def counter_per_dispatch(path, counter):
    """Average over dispatches of the counter (summed over its instances within a dispatch) per kernel."""
    per_dispatch = collections.defaultdict(float)
    names = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        per_dispatch[r["Dispatch_Id"]] += float(r["Counter_Value"])
        names[r["Dispatch_Id"]] = _kernel(r["Kernel_Name"])
    acc = collections.defaultdict(list)
    for d, v in per_dispatch.items():
        acc[names[d]].append(v)
    return {k: sum(v) / len(v) for k, v in acc.items()}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv")
    t.add_argument("write_csv")
    t.add_argument("--key", required=True)
    t.add_argument("--json")
    c = sub.add_parser("counters")
    c.add_argument("csvs", nargs="+")
    c.add_argument("--stats", required=True)
    for p in (s, t, c):
        p.add_argument("--title", required=True)
        p.add_argument("--command", action="append", default=[])
    a = ap.parse_args(argv)
    w = sys.stdout.write
    w(f"# {a.title}\n\nCommands (run from the repository root on an MI355X, `export TMPDIR=/tmp`):\n\n```\n")
    for c in a.command:
        w(c + "\n")
    w("```\n\n")
    if a.cmd == "stats":
        w("| kernel | calls | average us | min us | max us | % of GPU time |\n|---|---|---|---|---|---|\n")
        for name, calls, avg, mn, mx, pct in stats_table(a.csv):
            w(f"| `{name}` | {calls} | {avg:.1f} | {mn:.1f} | {mx:.1f} | {pct:.2f} |\n")
        return
    if a.cmd == "counters":
        names = []
        for path in a.csvs:
            names += sorted({r["Counter_Name"] for r in csv.DictReader(open(path))} - set(names))
        vals = {n: {} for n in names}
        for path in a.csvs:
            for n in {r["Counter_Name"] for r in csv.DictReader(open(path))}:
                vals[n].update(counter_per_dispatch(path, n))
        avg_us = {name: avg for name, _, avg, _, _, _ in stats_table(a.stats)}
        kernels = sorted({k for n in names for k in vals[n]}, key=lambda k: -avg_us.get(k, 0.))
        mfma = "SQ_INSTS_MFMA" in names
        w("Per-dispatch averages; `average us` from the kernel-trace statistics.\n\n")
        w("| kernel | average us | " + " | ".join(names) + (" | MFMA TFLOP/s | of 157.3 TF |" if mfma else " |") + "\n")
        w("|---|---|" + "---|" * len(names) + ("---|---|" if mfma else "") + "\n")
        for k in kernels:
            row = [f"`{k}`", f"{avg_us[k]:.1f}" if k in avg_us else "-"]
            row += [f"{vals[n][k]:.4g}" if k in vals[n] else "-" for n in names]
            if mfma:
                insts = vals["SQ_INSTS_MFMA"].get(k, 0.)
                if insts and k in avg_us:
                    tf = insts * MFMA_FLOP / (avg_us[k] * 1e-6) / 1e12
                    row += [f"{tf:.2f}", f"{tf * 1e12 / FP32_MATRIX_PEAK:.3f}"]
                else:
                    row += ["-", "-"]
            w("| " + " | ".join(row) + " |\n")
        return
    fetch = counter_per_dispatch(a.fetch_csv, "FETCH_SIZE")
    write = counter_per_dispatch(a.write_csv, "WRITE_SIZE")
    copy = [k for k in fetch if k.startswith(COPY_KERNEL)]
    if not copy:
        sys.exit("no hx_stream_copy dispatch in the profile: cannot calibrate")
    fs = COPY_BYTES / 1024 / fetch[copy[0]]
    ws = COPY_BYTES / 1024 / write[copy[0]]
    w(f"Calibration: `hx_stream_copy` moves {COPY_BYTES // 1024} KiB each way per dispatch and reads FETCH_SIZE = "
      f"{fetch[copy[0]]:.0f} KiB, WRITE_SIZE = {write[copy[0]]:.0f} KiB, so FETCH_SIZE is scaled by {fs:.4f} and "
      f"WRITE_SIZE by {ws:.4f} below.\n\n")
    w("| kernel | FETCH_SIZE KiB (raw) | WRITE_SIZE KiB (raw) | HBM bytes/launch (scaled) |\n|---|---|---|---|\n")
    table = {}
    for k in sorted(fetch, key=lambda k: -(fetch[k] * fs + write.get(k, 0) * ws)):
        b = 1024 * (fetch[k] * fs + write.get(k, 0) * ws)
        table[k] = b
        w(f"| `{k}` | {fetch[k]:.1f} | {write.get(k, 0):.1f} | {b:.4g} |\n")
    if a.json:
        data = json.load(open(a.json)) if os.path.exists(a.json) else {}
        data[a.key] = {"calibration": {"fetch_scale": fs, "write_scale": ws}, "bytes_per_launch": table}
        with open(a.json, "w") as f:
            json.dump(data, f, indent=1)
            f.write("\n")


if __name__ == "__main__":
    main()
