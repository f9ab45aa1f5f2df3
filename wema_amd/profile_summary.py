#!/usr/bin/env python3
"""Summarise rocprofv3 CSV output of a bench.py workload into a short markdown table and
a traffic file (HBM bytes per launch per kernel) that bench.py reads.

Inputs: one `--kernel-trace --stats` directory and one `--pmc FETCH_SIZE` and one
`--pmc WRITE_SIZE` directory of the same workload, each run with bench.py
--copy-calibration (the PMC runs), which adds a 1 GiB device copy: its counter readings
give the scale from counter units to bytes (FETCH_SIZE and WRITE_SIZE are in KiB; the copy
fetches and writes exactly 1 GiB). Optionally (--valu) a `--pmc SQ_INSTS_VALU SQ_INSTS_LDS
SQ_INSTS_SALU SQ_WAVES SQ_BUSY_CYCLES SQ_WAVE_CYCLES` directory: instructions per wave and
the VALU issue fraction, SQ_INSTS_VALU x 4 cycles (a wave64 instruction occupies a 16-lane
SIMD for 4 cycles) / (1024 SIMDs x 2.4 GHz x the kernel's mean duration in the trace),
written to valu_<workload>.json, which bench.py reports beside the HBM fraction.

    python -m wema_amd.profile_summary --workload c2 --title "..." \\
        --trace DIR --fetch DIR --write DIR --cmd "..." --out profiles
"""
import argparse
import collections
import csv
import glob
import json
import os
import re

GIB = float(1 << 30)


def _one(pattern):
    files = glob.glob(pattern)
    if len(files) != 1:
        raise SystemExit("expected one file for %s, found %d" % (pattern, len(files)))
    return files[0]


def _short(name):
    """Kernel name without its argument list; the guided filter's general kernel (k_gf_any)
    is reported under the name bench.py times both by (k_gf)."""
    n = name.split("(")[0]
    if n.startswith("void "):
        n = n[5:]
    return "k_gf" if n == "k_gf_any" else n


def pipeline_dispatches(rows):
    """Dispatch ids of the pipeline's own launches: from the first step (its level-0 forward
    DWT, the launch before its forward levels and temporal stage) on, which leaves out the
    autotuners' timing launches at session open (they run no temporal stage); device copies
    (the counter calibration) are always kept."""
    order = sorted({int(r["Dispatch_Id"]): _short(r["Kernel_Name"]) for r in rows}.items())
    names = [n for _, n in order]
    first_t = next((i for i, n in enumerate(names) if n in ("k_iir", "k_iir1") or n.startswith("k_win")), None)
    cut = 0
    if first_t is not None:
        cut = first_t
        while cut > 0 and names[cut - 1].startswith("k_dwt_fwd"):
            cut -= 1
            if names[cut] == "k_dwt_fwd0":
                break
    return {d for i, (d, n) in enumerate(order) if i >= cut or n.startswith("__amd_rocclr_copy")}


def kernel_stats(trace_dir):
    """Per (kernel, grid size): calls, mean / min / max duration in us, share of GPU time.
    The DWT kernels run one grid size per level, so each level gets its own row."""
    per = collections.defaultdict(list)
    rows = list(csv.DictReader(open(_one(os.path.join(trace_dir, "*_kernel_trace.csv")))))
    keep = pipeline_dispatches(rows)
    for r in rows:
        if int(r["Dispatch_Id"]) not in keep:
            continue
        grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        per[(_short(r["Kernel_Name"]), grid)].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    total = sum(sum(v) for v in per.values())
    rows = [(n, g, len(v), sum(v) / len(v), min(v), max(v), 100.0 * sum(v) / total) for (n, g), v in per.items()]
    return sorted(rows, key=lambda r: -r[3] * r[2])


def counters(pmc_dir, counter):
    rows = list(csv.DictReader(open(_one(os.path.join(pmc_dir, "*_counter_collection.csv")))))
    keep = pipeline_dispatches(rows)
    per = collections.defaultdict(list)
    for r in rows:
        if r["Counter_Name"] == counter and int(r["Dispatch_Id"]) in keep:
            per[_short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return per


SIMDS = 1024          # 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9      # peak engine clock
VALU_COUNTERS = ("SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_INSTS_SALU", "SQ_WAVES", "SQ_BUSY_CYCLES",
                 "SQ_WAVE_CYCLES")


def valu_stats(pmc_dir, stats):
    """Per kernel: mean counter values per launch, instructions per wave, VALU issue fraction."""
    per = {c: counters(pmc_dir, c) for c in VALU_COUNTERS}
    dur = collections.defaultdict(list)
    for n, g, calls, avg, mn, mx, pct in stats:
        dur[n] += [avg] * calls
    out = {}
    for name in sorted(per["SQ_INSTS_VALU"]):
        if not name.startswith("k_") or name == "k_synth" or not dur.get(name):
            continue
        mean = {c: sum(per[c].get(name, [0])) / max(len(per[c].get(name, [])), 1) for c in VALU_COUNTERS}
        waves = max(mean["SQ_WAVES"], 1.0)
        us = sum(dur[name]) / len(dur[name])
        out[name] = {"launches": len(per["SQ_INSTS_VALU"][name]), "mean_us": round(us, 2),
                     "waves": round(waves), "valu_per_wave": round(mean["SQ_INSTS_VALU"] / waves, 1),
                     "lds_per_wave": round(mean["SQ_INSTS_LDS"] / waves, 1),
                     "salu_per_wave": round(mean["SQ_INSTS_SALU"] / waves, 1),
                     "valu_frac": round(mean["SQ_INSTS_VALU"] * 4.0 / (SIMDS * CLOCK_HZ * us * 1e-6), 4)}
    return out


def calibration(per, counter):
    """Bytes per counter unit, from the 1 GiB copy: the copy dispatches are the device
    copies whose reading is the largest (the other copies are small table uploads)."""
    copies = per.get("__amd_rocclr_copyBuffer", [])
    if not copies:
        raise SystemExit("no copy kernel in the %s run (bench.py --copy-calibration)" % counter)
    big = max(copies)
    return GIB / (big * 1024.0), big


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", required=True)
    ap.add_argument("--title", required=True)
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--cmd", action="append", default=[], help="command lines that made the data")
    ap.add_argument("--out", default="profiles")
    ap.add_argument("--md-name", required=True)
    ap.add_argument("--valu", help="--pmc run of the SQ instruction counters (optional)")
    ap.add_argument("--pixels-per-launch", type=float, default=0.0,
                    help="output pixels per launch, for VALU instructions per pixel")
    a = ap.parse_args()

    stats = kernel_stats(a.trace)
    fetch = counters(a.fetch, "FETCH_SIZE")
    write = counters(a.write, "WRITE_SIZE")
    fs, fraw = calibration(fetch, "FETCH_SIZE")
    ws, wraw = calibration(write, "WRITE_SIZE")

    traffic = {}
    for name in sorted(set(fetch) | set(write)):
        if not name.startswith("k_") or name == "k_synth":
            continue
        fb = sum(fetch.get(name, [0])) / max(len(fetch.get(name, [])), 1) * 1024.0 * fs
        wb = sum(write.get(name, [0])) / max(len(write.get(name, [])), 1) * 1024.0 * ws
        traffic[name] = {"fetch_bytes_per_launch": round(fb), "write_bytes_per_launch": round(wb),
                         "hbm_bytes_per_launch": round(fb + wb), "launches": len(fetch.get(name, []))}
    os.makedirs(a.out, exist_ok=True)
    with open(os.path.join(a.out, "traffic_%s.json" % a.workload), "w") as f:
        json.dump({"workload": a.workload, "fetch_scale": round(fs, 4), "write_scale": round(ws, 4),
                   "kernels": traffic}, f, indent=1)
        f.write("\n")

    lines = ["# %s" % a.title, "", "Commands (from the repository root on the GPU host):", ""]
    lines += ["    " + c for c in a.cmd]
    lines += ["", "## Kernel durations (rocprofv3 --kernel-trace --stats)", "",
              "| kernel | grid (threads) | calls | avg us | min us | max us | % of GPU time |",
              "|---|---:|---:|---:|---:|---:|---:|"]
    for n, g, c, avg, mn, mx, pct in stats:
        lines.append("| %s | %d | %d | %.1f | %.1f | %.1f | %.2f |" % (n, g, c, avg, mn, mx, pct))
    lines += ["", "## HBM traffic per launch (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE)", "",
              "Scale from the 1 GiB device copy of the same runs: FETCH_SIZE read %.0f KiB "
              "(x%.3f), WRITE_SIZE read %.0f KiB (x%.3f)." % (fraw, fs, wraw, ws), "",
              "| kernel | launches | fetch MB | write MB | HBM MB per launch |", "|---|---:|---:|---:|---:|"]
    for n, t in sorted(traffic.items(), key=lambda kv: -kv[1]["hbm_bytes_per_launch"]):
        lines.append("| %s | %d | %.1f | %.1f | %.1f |" % (
            n, t["launches"], t["fetch_bytes_per_launch"] / 1e6, t["write_bytes_per_launch"] / 1e6,
            t["hbm_bytes_per_launch"] / 1e6))
    if a.valu:
        vs = valu_stats(a.valu, stats)
        with open(os.path.join(a.out, "valu_%s.json" % a.workload), "w") as f:
            json.dump({"workload": a.workload, "simds": SIMDS, "clock_hz": CLOCK_HZ, "kernels": vs}, f, indent=1)
            f.write("\n")
        lines += ["", "## Instructions (rocprofv3 --pmc %s)" % " ".join(VALU_COUNTERS), "",
                  "VALU issue fraction = SQ_INSTS_VALU x 4 cycles / (%d SIMDs x %.1f GHz x mean duration)."
                  % (SIMDS, CLOCK_HZ / 1e9), "",
                  "| kernel | launches | waves per launch | VALU / wave | LDS / wave | SALU / wave |%s VALU issue fraction |"
                  % (" VALU lane-ops / pixel |" if a.pixels_per_launch else ""),
                  "|---|---:|---:|---:|---:|---:|%s---:|" % ("---:|" if a.pixels_per_launch else "")]
        for n, v in sorted(vs.items(), key=lambda kv: -kv[1]["mean_us"]):
            ppx = ""
            if a.pixels_per_launch:
                ppx = " %.1f |" % (v["valu_per_wave"] * v["waves"] * 64 / a.pixels_per_launch)
            lines.append("| %s | %d | %d | %.1f | %.1f | %.1f |%s %.3f |" % (
                n, v["launches"], v["waves"], v["valu_per_wave"], v["lds_per_wave"], v["salu_per_wave"],
                ppx, v["valu_frac"]))
    lines.append("")
    with open(os.path.join(a.out, a.md_name), "w") as f:
        f.write("\n".join(lines))
    print("wrote", os.path.join(a.out, a.md_name), "and traffic_%s.json" % a.workload)


if __name__ == "__main__":
    main()
