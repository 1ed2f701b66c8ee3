"""Summarise rocprofv3 CSV output into the short text records kept under profiles/.

    python -m sfsim_amd.profile_summary stats <kernel_stats.csv> <out.md> <command...>
    python -m sfsim_amd.profile_summary pmc <counter_collection.csv> <kernel-regex>
    python -m sfsim_amd.profile_summary traffic <fetch.csv> <write.csv> <calib_fetch.csv> <calib_write.csv> \
        <calib_mib> <kernel-regex> <out.json> <command...>
"""
import csv
import os
import statistics
import json
import re
import sys


def _short(name):
    m = re.search(r"(\w+_kernel|__amd_rocclr_\w+|\w+(?=\())", name)
    return m.group(1) if m else name[:40]


def stats(path, out, command):
    """Kernel statistics table; with the kernel trace next to the stats file (k_kernel_trace.csv for
    k_kernel_stats.csv) also each kernel's median duration, which the few cold first launches do not move."""
    rows = list(csv.DictReader(open(path)))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    medians = {}
    trace = path.replace("_kernel_stats.csv", "_kernel_trace.csv")
    if trace != path and os.path.exists(trace):
        durations = {}
        for t in csv.DictReader(open(trace)):
            durations.setdefault(t["Kernel_Name"], []).append(int(t["End_Timestamp"]) - int(t["Start_Timestamp"]))
        medians = {k: statistics.median(v) for k, v in durations.items()}
    lines = [f"# rocprofv3 kernel trace: {command}", "",
             "| kernel | calls | total ms | average us | median us | min us | max us | % |",
             "|---|---|---|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        med = medians.get(r["Name"])
        med_s = f"{med / 1e3:.1f}" if med is not None else "-"
        lines.append(f"| {_short(r['Name'])} | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                     f"{float(r['AverageNs']) / 1e3:.1f} | {med_s} | {float(r['MinNs']) / 1e3:.1f} | "
                     f"{float(r['MaxNs']) / 1e3:.1f} | {100 * float(r['TotalDurationNs']) / total:.1f} |")
    open(out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def per_dispatch(path, regex, counter):
    """Average counter value per dispatch of kernels whose name matches regex."""
    vals = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter or not re.search(regex, r["Kernel_Name"]):
            continue
        vals[r["Dispatch_Id"]] = vals.get(r["Dispatch_Id"], 0.0) + float(r["Counter_Value"])
    if not vals:
        raise SystemExit(f"no {counter} rows for {regex} in {path}")
    return sum(vals.values()) / len(vals), len(vals)


def traffic(fetch_csv, write_csv, cal_fetch_csv, cal_write_csv, cal_mib, regex, out, command):
    cal_bytes = float(cal_mib) * (1 << 20)
    cf, _ = per_dispatch(cal_fetch_csv, "copy|Copy|elementwise|vectorized", "FETCH_SIZE")
    cw, _ = per_dispatch(cal_write_csv, "copy|Copy|elementwise|vectorized", "WRITE_SIZE")
    f_scale = cal_bytes / (cf * 1024.0)
    w_scale = cal_bytes / (cw * 1024.0)
    fk, nf = per_dispatch(fetch_csv, regex, "FETCH_SIZE")
    wk, nw = per_dispatch(write_csv, regex, "WRITE_SIZE")
    rec = {"kernel": regex, "command": command, "fetch_kib_per_dispatch": fk, "write_kib_per_dispatch": wk,
           "dispatches": [nf, nw], "calibration_mib": cal_mib, "fetch_scale": f_scale, "write_scale": w_scale,
           "bytes_per_launch": fk * 1024.0 * f_scale + wk * 1024.0 * w_scale}
    json.dump(rec, open(out, "w"), indent=1)
    print(json.dumps(rec, indent=1))


if __name__ == "__main__":
    mode = sys.argv[1]
    if mode == "stats":
        stats(sys.argv[2], sys.argv[3], " ".join(sys.argv[4:]))
    elif mode == "pmc":
        print(per_dispatch(sys.argv[2], sys.argv[3], sys.argv[4]))
    elif mode == "traffic":
        traffic(*sys.argv[2:9], " ".join(sys.argv[9:]))
