"""Summaries of rocprofv3 output for the profiles/ folder.

    python -m mujoco_contact_surfaces_amd.prof_summary kernels DIR [--only REGEX] [--command CMD]
    python -m mujoco_contact_surfaces_amd.prof_summary counters --kernels REGEX \
        [--fetch DIR --write DIR --calib-fetch DIR --calib-write DIR --calib-mib N] [--pass DIR ...] \
        [--cus 256] [--json OUT] [--command CMD ...]

`kernels` reads the *kernel_stats.csv and *kernel_trace.csv of one
`rocprofv3 --kernel-trace --stats --output-format csv` run and prints a
markdown table per kernel (calls, total, average, min, max) and the average
duration per dispatch shape (grid size).

`counters` reads *counter_collection.csv files of `rocprofv3 --pmc` runs (one
directory per run; rocprofv3 collects each run's counters in one pass) and
prints, for the last dispatch of every kernel matching --kernels:
  * HBM bytes: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by
    what the same counter reads for a device-to-device copy of a known size
    (bench.py --copy-calibration N) -- bytes = counter x 1024 x (copy bytes /
    (counter of the copy x 1024));
  * occupancy: SQ_WAVE_CYCLES / GRBM_GUI_ACTIVE / CUs (mean resident waves per
    CU; on gfx950 SQ_WAVE_CYCLES counts in units of 4 cycles, the factor the
    OccupancyPercent expression in rocprofiler-sdk's counter_defs.yaml uses),
    and SQ_WAVES (waves launched).  GRBM_GUI_ACTIVE comes summed over the
    accelerator complex dies (XCDs), so it is divided by their number
    (Num_Xcc of the run's agent_info.csv) to give the dispatch's clocks;
  * vector-memory path: TD_TD_BUSY_sum and TA_TA_BUSY_sum over GRBM_GUI_ACTIVE
    and CUs (busy fraction of the texture data / address units), the mean L2
    read latency TCP_TCC_READ_REQ_LATENCY_sum / TCP_TCC_READ_REQ_sum (cycles);
  * vector ALU: SQ_INSTS_VALU over GRBM_GUI_ACTIVE and CUs (issue use: a
    CU's four SIMDs issue one wave64 instruction each per four cycles, so one
    per cycle per CU at most), SQ_THREAD_CYCLES_VALU / (64 SQ_INSTS_VALU)
    (active lanes per instruction), SQ_WAIT_ANY / SQ_WAVE_CYCLES (share of
    the waves' time spent waiting).
Counters a run did not collect are left out.  Nothing here runs on the GPU.
"""
import argparse
import csv
import glob
import json
import os
import re
import sys
from collections import OrderedDict, defaultdict


def _files(d, suffix):
    return sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))


def _rows(d, suffix):
    out = []
    for f in _files(d, suffix):
        with open(f, newline="") as fh:
            out.extend(csv.DictReader(fh))
    return out


def _col(row, *names):
    for n in names:
        if n in row and row[n] != "":
            return row[n]
    raise KeyError(names)


def _short(name, width=90):
    name = re.sub(r"\s+", " ", name)
    return name if len(name) <= width else name[:width - 3] + "..."


def kernels(args):
    stats = _rows(args.dir, "kernel_stats.csv")
    trace = _rows(args.dir, "kernel_trace.csv")
    if not stats:
        sys.exit("no kernel_stats.csv under %s" % args.dir)
    only = re.compile(args.only) if args.only else None
    lines = []
    if args.command:
        lines += ["Command (repository root, one MI355X, ROCm 7.2):", "", "    " + args.command, ""]
    lines += ["| kernel | calls | total ms | avg us | min us | max us | % of traced time |",
              "|---|---|---|---|---|---|---|"]
    for r in stats:
        name = _col(r, "Name", "KernelName", "Kernel_Name")
        if only and not only.search(name):
            continue
        lines.append("| `%s` | %s | %.3f | %.1f | %.1f | %.1f | %.2f |" % (
            _short(name), _col(r, "Calls"), float(_col(r, "TotalDurationNs")) / 1e6,
            float(_col(r, "AverageNs")) / 1e3, float(_col(r, "MinNs")) / 1e3, float(_col(r, "MaxNs")) / 1e3,
            float(_col(r, "Percentage"))))
    if trace:
        shapes = defaultdict(list)
        for r in trace:
            name = _col(r, "Kernel_Name", "KernelName", "Name")
            if only and not only.search(name):
                continue
            grid = _col(r, "Grid_Size_X", "Grid_Size", "GridSizeX")
            dur = (int(_col(r, "End_Timestamp")) - int(_col(r, "Start_Timestamp"))) / 1e3
            shapes[(name, grid)].append(dur)
        lines += ["", "Per dispatch shape (grid x in threads):", "",
                  "| kernel | grid x | dispatches | avg us | min us | max us |", "|---|---|---|---|---|---|"]
        for (name, grid), ds in sorted(shapes.items(), key=lambda kv: -sum(kv[1])):
            lines.append("| `%s` | %s | %d | %.1f | %.1f | %.1f |" % (_short(name), grid, len(ds),
                                                                      sum(ds) / len(ds), min(ds), max(ds)))
    print("\n".join(lines))


def _last_dispatch_values(d, kernel_re):
    """{kernel: {counter: value}} for the last dispatch of each kernel."""
    rows = _rows(d, "counter_collection.csv")
    per = defaultdict(lambda: defaultdict(dict))
    for r in rows:
        name = _col(r, "Kernel_Name", "KernelName")
        if not kernel_re.search(name):
            continue
        disp = int(_col(r, "Dispatch_Id", "Correlation_Id"))
        per[name][disp][_col(r, "Counter_Name")] = float(_col(r, "Counter_Value"))
    out = {}
    for name, by in per.items():
        out[name] = by[max(by)]
    return out


def _num_xcc(dirs):
    for d in dirs:
        for r in _rows(d, "agent_info.csv"):
            if r.get("Agent_Type") == "GPU" and r.get("Num_Xcc"):
                return int(r["Num_Xcc"])
    return 1


def _copy_counter(d, counter):
    rows = _rows(d, "counter_collection.csv")
    vals = [float(_col(r, "Counter_Value")) for r in rows
            if _col(r, "Counter_Name") == counter and "copy" in _col(r, "Kernel_Name", "KernelName").lower()]
    if not vals:
        sys.exit("no copy kernel with %s under %s" % (counter, d))
    return max(vals)


def counters(args):
    kre = re.compile(args.kernels)
    table = OrderedDict()

    def merge(vals):
        for k, v in vals.items():
            table.setdefault(k, {}).update(v)

    for d in [args.fetch, args.write] + (args.passes or []):
        if d:
            merge(_last_dispatch_values(d, kre))
    scale = {}
    copy_bytes = args.calib_mib << 20
    if args.fetch and args.calib_fetch:
        scale["FETCH_SIZE"] = copy_bytes / (_copy_counter(args.calib_fetch, "FETCH_SIZE") * 1024.0)
    if args.write and args.calib_write:
        scale["WRITE_SIZE"] = copy_bytes / (_copy_counter(args.calib_write, "WRITE_SIZE") * 1024.0)
    lines = []
    for c in args.command or []:
        lines.append("    " + c)
    if lines:
        lines = ["Commands (repository root, one MI355X, ROCm 7.2), one counter pass per run:", ""] + lines + [""]
    if scale:
        lines.append("Calibration: a device-to-device copy of %d MiB; bytes = counter x 1024 x %s." % (
            args.calib_mib, ", ".join("%.3f (%s)" % (v, k) for k, v in scale.items())))
        lines.append("")
    cols = ["FETCH_SIZE", "WRITE_SIZE", "SQ_WAVES", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "GRBM_GUI_ACTIVE",
            "TD_TD_BUSY_sum", "TA_TA_BUSY_sum", "TCP_TCC_READ_REQ_LATENCY_sum", "TCP_TCC_READ_REQ_sum",
            "SQ_INSTS_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_WAIT_ANY"]
    cols = [c for c in cols if any(c in v for v in table.values())]
    lines.append("Last dispatch of each kernel:")
    lines.append("")
    lines.append("| kernel | " + " | ".join(c for c in cols) + " |")
    lines.append("|---|" + "---|" * len(cols))
    derived = OrderedDict()
    xcc = args.xcc or _num_xcc([d for d in [args.fetch, args.write] + (args.passes or []) if d])
    for k, v in table.items():
        lines.append("| `%s` | " % _short(k, 50) + " | ".join("%.0f" % v[c] if c in v else "" for c in cols) + " |")
        dv = OrderedDict()
        if "FETCH_SIZE" in v and "FETCH_SIZE" in scale:
            dv["hbm_read_bytes"] = v["FETCH_SIZE"] * 1024.0 * scale["FETCH_SIZE"]
        if "WRITE_SIZE" in v and "WRITE_SIZE" in scale:
            dv["hbm_write_bytes"] = v["WRITE_SIZE"] * 1024.0 * scale["WRITE_SIZE"]
        if "hbm_read_bytes" in dv and "hbm_write_bytes" in dv:
            dv["hbm_bytes"] = dv["hbm_read_bytes"] + dv["hbm_write_bytes"]
        g = v.get("GRBM_GUI_ACTIVE")
        if g:
            g = g / xcc
            dv["gpu_clocks"] = g
            if "SQ_WAVE_CYCLES" in v:
                dv["mean_waves_per_cu"] = 4.0 * v["SQ_WAVE_CYCLES"] / g / args.cus
                dv["occupancy_pct_of_32"] = 100.0 * dv["mean_waves_per_cu"] / 32.0
            if "TD_TD_BUSY_sum" in v:
                dv["td_busy_pct"] = 100.0 * v["TD_TD_BUSY_sum"] / g / args.cus
            if "TA_TA_BUSY_sum" in v:
                dv["ta_busy_pct"] = 100.0 * v["TA_TA_BUSY_sum"] / g / args.cus
            if "SQ_INSTS_VALU" in v:
                dv["valu_issue_pct"] = 100.0 * v["SQ_INSTS_VALU"] / g / args.cus
        if v.get("SQ_INSTS_VALU") and "SQ_THREAD_CYCLES_VALU" in v:
            dv["valu_lanes_pct"] = 100.0 * v["SQ_THREAD_CYCLES_VALU"] / (64.0 * v["SQ_INSTS_VALU"])
        if v.get("SQ_WAVE_CYCLES") and "SQ_WAIT_ANY" in v:
            dv["wait_pct"] = 100.0 * v["SQ_WAIT_ANY"] / v["SQ_WAVE_CYCLES"]
        if v.get("TCP_TCC_READ_REQ_sum"):
            dv["l2_read_latency_cycles"] = v.get("TCP_TCC_READ_REQ_LATENCY_sum", 0.0) / v["TCP_TCC_READ_REQ_sum"]
        if "SQ_WAVES" in v:
            dv["waves"] = v["SQ_WAVES"]
        derived[k] = dv
    lines += ["", "Derived (GRBM_GUI_ACTIVE / %d XCDs is the dispatch's duration in GPU clocks; %d CUs):" % (xcc, args.cus),
              "",
              "| kernel | HBM read GB | HBM write GB | HBM GB | mean waves / CU | occupancy % of 32 | TD busy % | "
              "TA busy % | L2 read latency (cycles) | VALU issue % | VALU lanes % | waiting % |",
              "|---|---|---|---|---|---|---|---|---|---|---|---|"]

    def f(x, fmt):
        return fmt % x if x is not None else ""

    for k, dv in derived.items():
        lines.append("| `%s` | %s | %s | %s | %s | %s | %s | %s | %s | %s | %s | %s |" % (
            _short(k, 50), f(dv.get("hbm_read_bytes", None) and dv["hbm_read_bytes"] / 1e9, "%.3f"),
            f(dv.get("hbm_write_bytes", None) and dv["hbm_write_bytes"] / 1e9, "%.3f"),
            f(dv.get("hbm_bytes", None) and dv["hbm_bytes"] / 1e9, "%.3f"), f(dv.get("mean_waves_per_cu"), "%.2f"),
            f(dv.get("occupancy_pct_of_32"), "%.1f"), f(dv.get("td_busy_pct"), "%.1f"),
            f(dv.get("ta_busy_pct"), "%.1f"), f(dv.get("l2_read_latency_cycles"), "%.0f"),
            f(dv.get("valu_issue_pct"), "%.1f"), f(dv.get("valu_lanes_pct"), "%.1f"), f(dv.get("wait_pct"), "%.1f")))
    print("\n".join(lines))
    if args.json:
        short = {}
        for k, dv in derived.items():
            m = re.search(r"(k_[a-z_0-9]+)", k)
            short[m.group(1) if m else k] = dv
        with open(args.json, "w") as fh:
            json.dump(dict(source=args.source, workload=args.workload, num_xcc=xcc, per_kernel=short,
                           hbm_bytes_per_launch={k: v["hbm_bytes"] for k, v in short.items() if "hbm_bytes" in v}),
                      fh, indent=1)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="mode", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("dir")
    k.add_argument("--only", default="")
    k.add_argument("--command", default="")
    c = sub.add_parser("counters")
    c.add_argument("--kernels", required=True)
    c.add_argument("--fetch")
    c.add_argument("--write")
    c.add_argument("--calib-fetch")
    c.add_argument("--calib-write")
    c.add_argument("--calib-mib", type=int, default=1024)
    c.add_argument("--pass", dest="passes", action="append")
    c.add_argument("--cus", type=int, default=256)
    c.add_argument("--xcc", type=int, default=0, help="XCDs (default: Num_Xcc of the runs' agent_info.csv, else 1)")
    c.add_argument("--json")
    c.add_argument("--source", default="")
    c.add_argument("--workload", default="")
    c.add_argument("--command", action="append")
    a = ap.parse_args(argv)
    (kernels if a.mode == "kernels" else counters)(a)


if __name__ == "__main__":
    main()
