"""Profiling records: one bench.py step joined with rocprofv3 output.

The reference brackets its kernel with bsg_cuda_print_stat_kernel_start/end
(apps/sgemm/kernel.cpp:288,329), which write a "vanilla_stats" region-of-
interest record, and each app ships a parser that turns that record into
numbers (apps/bfs/parse.py:12-27).  Here the region of interest is one step of
bench.py, and the record joins

  * the bench JSON line (workload, flops and algorithmic bytes per step, the
    roofline peak, the live HIP-event timing),
  * `rocprofv3 --kernel-trace --stats` (kernel_stats.csv: per-kernel calls and
    durations),
  * `rocprofv3 --pmc ...` runs (counter_collection.csv: per-dispatch counters),
  * the FETCH_SIZE / WRITE_SIZE readings of a device copy of known size, which
    scale the byte counters to bytes,

into one JSON object of schema SCHEMA, and renders it as the Markdown summary
kept under profiles/.  Command line:

  python -m hb_hammerbench_amd.profile_records --bench step.json \\
      --stats OUT/fprof_fp32 --pmc OUT/fpmc_fp32_FETCH_SIZE OUT/fpmc_fp32_WRITE_SIZE \\
      OUT/fpmc_fp32_SQ --copy-fetch OUT/fpmc_copy_FETCH_SIZE --copy-write OUT/fpmc_copy_WRITE_SIZE \\
      --dominant sgemm_f32_kernel --json rec.json --markdown rec.md --title "..." --command "..."

Only the standard library is used.
"""
import argparse
import csv
import glob
import json
import os
import sys

SCHEMA = "hbsgemm-profile/1"

# MI355X: 256 CUs = 1024 SIMDs in 8 XCDs.  GRBM_GUI_ACTIVE is summed over the
# XCDs, so the clock during a dispatch is GRBM_GUI_ACTIVE / XCDS / duration;
# SQ_VALU_MFMA_BUSY_CYCLES is summed over the SIMDs.
XCDS = 8
SIMDS = 1024
# matrix-core operations per SQ_INSTS_VALU_MFMA_MOPS_* unit
MOPS_UNIT = 512
MOPS_COUNTERS = ("SQ_INSTS_VALU_MFMA_MOPS_F32", "SQ_INSTS_VALU_MFMA_MOPS_BF16",
                 "SQ_INSTS_VALU_MFMA_MOPS_F8")
# launches that prepare a benchmark's inputs, not part of a timed step
HELPER_KERNELS = ("fill_random", "fill_pattern", "__amd_rocclr")
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def short_name(name):
    """Kernel name without its argument list, `void ` and anonymous namespace."""
    n = name.replace("void ", "").replace("(anonymous namespace)::", "")
    return n.split("((")[0] if "((" in n else n.split("(")[0]


def find_csv(path, suffix):
    """`path` itself if it is a file, else the one file under it ending in `suffix`."""
    if os.path.isfile(path):
        return path
    hits = sorted(glob.glob(os.path.join(path, "**", "*" + suffix), recursive=True))
    if not hits:
        raise FileNotFoundError(f"no *{suffix} under {path}")
    return hits[0]


def read_kernel_stats(path):
    """Rows of a rocprofv3 kernel_stats.csv as dicts with numeric fields."""
    rows = []
    with open(find_csv(path, "kernel_stats.csv"), newline="") as f:
        for r in csv.DictReader(f):
            rows.append({"name": short_name(r["Name"]), "calls": int(r["Calls"]),
                         "total_ns": float(r["TotalDurationNs"]), "avg_ns": float(r["AverageNs"]),
                         "min_ns": float(r["MinNs"]), "max_ns": float(r["MaxNs"]),
                         "percent": float(r["Percentage"])})
    return rows


def read_kernel_trace(path):
    """{kernel: [duration ns of each dispatch, in dispatch order]} from a
    rocprofv3 kernel_trace.csv (the --kernel-trace output)."""
    out = {}
    rows = []
    with open(find_csv(path, "kernel_trace.csv"), newline="") as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), short_name(r["Kernel_Name"]),
                         int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))
    for _, name, dur in sorted(rows):
        out.setdefault(name, []).append(dur)
    return out


def skip_warmup(stats, trace, warmup):
    """Kernel statistics with each kernel's first `warmup` dispatches left out
    (the untimed warm-up launches of a bench run), so that averages cover the
    same launches as the bench line's HIP-event timing.  Kernels launched no
    more than `warmup` times keep their figures."""
    out = []
    for r in stats:
        d = trace.get(r["name"], [])
        if warmup > 0 and len(d) > warmup:
            t = d[warmup:]
            r = dict(r, calls=len(t), total_ns=float(sum(t)), avg_ns=sum(t) / len(t),
                     min_ns=float(min(t)), max_ns=float(max(t)), warmup_skipped=warmup)
        out.append(r)
    return out


def read_counters(*paths):
    """{(kernel, counter): mean value per dispatch} over rocprofv3
    counter_collection.csv files (one per --pmc run)."""
    acc = {}
    for p in paths:
        with open(find_csv(p, "counter_collection.csv"), newline="") as f:
            for r in csv.DictReader(f):
                key = (short_name(r["Kernel_Name"]), r["Counter_Name"])
                acc.setdefault(key, []).append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def copy_scales(copy_fetch, copy_write, copy_bytes=1 << 30):
    """Bytes per counted byte for FETCH_SIZE and WRITE_SIZE, from the copy
    kernel's readings (KiB per dispatch) for a copy of `copy_bytes` each way."""
    kib = copy_bytes / 1024.0
    fetch = [v for (k, c), v in copy_fetch.items() if COPY_KERNEL in k and c == "FETCH_SIZE"]
    write = [v for (k, c), v in copy_write.items() if COPY_KERNEL in k and c == "WRITE_SIZE"]
    if not fetch or not write:
        raise ValueError("copy calibration: no FETCH_SIZE / WRITE_SIZE for " + COPY_KERNEL)
    return {"fetch_kib": fetch[0], "write_kib": write[0], "copy_bytes": copy_bytes,
            "scale_fetch": kib / fetch[0], "scale_write": kib / write[0]}


def make_record(bench, stats, counters, dominant=None, scales=None, commands=()):
    """The profiling record of one bench step.

    bench     the bench.py JSON line (dict)
    stats     read_kernel_stats() rows of the --kernel-trace --stats run
    counters  read_counters() of the --pmc runs
    dominant  substring of the dominant kernel's name (default: the step
              kernel with the largest total duration)
    scales    copy_scales() result, or None (then no byte figures)
    """
    cfg = bench["config"]
    flops = float(cfg["flops_per_step"])
    peak = float(bench["roofline"]["peak"])
    step = [r for r in stats if not any(h in r["name"] for h in HELPER_KERNELS)]
    if not step:
        raise ValueError("no step kernel in the kernel statistics")
    if dominant:
        match = [r for r in step if dominant in r["name"]]
        if not match:
            raise ValueError(f"dominant kernel {dominant!r} not among {[r['name'] for r in step]}")
        dom = match[0]
    else:
        dom = max(step, key=lambda r: r["total_ns"])
    names = {r["name"] for r in step}
    dom_s = dom["avg_ns"] * 1e-9
    step_s = sum(r["avg_ns"] for r in step) * 1e-9  # each step kernel launches once per step

    def ctr(counter, kernel=dom["name"]):
        return counters.get((kernel, counter))

    grbm = ctr("GRBM_GUI_ACTIVE")
    busy = ctr("SQ_VALU_MFMA_BUSY_CYCLES")
    mops = {c: ctr(c) for c in MOPS_COUNTERS if ctr(c) is not None}
    derived = {
        "tflops_dominant": flops / dom_s / 1e12,
        "frac_dominant": flops / dom_s / 1e12 / peak,
        "tflops_step": flops / step_s / 1e12,
        "clock_ghz": grbm / XCDS / dom_s / 1e9 if grbm else None,
        "mfma_busy": busy / (SIMDS * grbm / XCDS) if (grbm and busy is not None) else None,
        "matrix_ops_executed": sum(mops.values()) * MOPS_UNIT if mops else None,
    }
    traffic = None
    fetch = [v for (k, c), v in counters.items() if k in names and c == "FETCH_SIZE"]
    write = [v for (k, c), v in counters.items() if k in names and c == "WRITE_SIZE"]
    if scales and fetch and write:
        fe, wr = sum(fetch), sum(write)
        traffic = {"fetch_kib": fe, "write_kib": wr, "scale_fetch": scales["scale_fetch"],
                   "scale_write": scales["scale_write"],
                   "hbm_bytes": (fe * scales["scale_fetch"] + wr * scales["scale_write"]) * 1024.0}
        if cfg.get("bytes_algorithmic"):
            traffic["bytes_algorithmic"] = float(cfg["bytes_algorithmic"])
    return {
        "schema": SCHEMA,
        "workload": cfg.get("workload"),
        "variant": cfg.get("variant"),
        "flops_per_step": flops,
        "bytes_algorithmic": cfg.get("bytes_algorithmic"),
        "peak_tflops": peak,
        "bench": {"value": bench.get("value"), "unit": bench.get("unit"),
                  "kernel_ms_avg": bench["roofline"].get("kernel_ms_avg"),
                  "frac": bench["roofline"].get("frac")},
        "kernels": stats,
        "step_kernels": sorted(names),
        "dominant": {"name": dom["name"], "avg_ns": dom["avg_ns"],
                     "counters": {c: v for (k, c), v in counters.items() if k == dom["name"]}},
        "derived": derived,
        "traffic": traffic,
        "commands": list(commands),
    }


def to_markdown(rec, title):
    """The Markdown summary of a record (the profiles/*.md layout)."""
    d, dom = rec["derived"], rec["dominant"]
    lines = [f"# {title}", ""]
    if rec["commands"]:
        lines += ["Commands (`$REPO` = repository root, one MI355X):", "", "```"]
        lines += rec["commands"] + ["```", ""]
    skipped = max((r.get("warmup_skipped", 0) for r in rec["kernels"]), default=0)
    head = "## Kernel stats (`kernel_stats.csv`, one step = the launches below)"
    if skipped:
        head = (f"## Kernel stats (`kernel_trace.csv` without each kernel's first {skipped} "
                "dispatches, the bench's warm-up; one step = the launches below)")
    lines += [head, "",
              "| kernel | calls | avg ns | min ns | max ns | % |", "|---|---|---|---|---|---|"]
    for r in rec["kernels"]:
        lines.append(f"| `{r['name']}` | {r['calls']} | {r['avg_ns']:,.0f} | {r['min_ns']:,.0f} | "
                     f"{r['max_ns']:,.0f} | {r['percent']:.2f} |")
    lines += ["", f"## Counters per dispatch of `{dom['name']}`", "", "| counter | value |",
              "|---|---|"]
    for c in sorted(dom["counters"]):
        if c not in ("FETCH_SIZE", "WRITE_SIZE"):
            lines.append(f"| `{c}` | {dom['counters'][c]:,.0f} |")
    t = rec["traffic"]
    if t:
        lines += [f"| `FETCH_SIZE` (KiB, all launches of a step) | {t['fetch_kib']:,.0f} |",
                  f"| `WRITE_SIZE` (KiB, all launches of a step) | {t['write_kib']:,.0f} |"]
    lines += ["", "## Derived", ""]
    work = f"* algorithmic work per step: {rec['flops_per_step']:.4e} FP32-equivalent flops"
    if rec.get("bytes_algorithmic"):
        work += f", {rec['bytes_algorithmic'] / 1e9:.2f} GB algorithmic bytes"
    lines.append(work)
    if d["matrix_ops_executed"] is not None:
        lines.append("* matrix-core ops executed (MOPS x 512, all precisions, incl. tile padding): "
                     f"{d['matrix_ops_executed']:.4e}")
    lines.append(f"* {dom['name']}: {dom['avg_ns'] / 1e6:.3f} ms -> **{d['tflops_dominant']:.1f} "
                 f"TFLOP/s** = {d['frac_dominant'] * 100:.1f}% of {rec['peak_tflops']:.1f}")
    lines.append(f"* whole step (all launches): {rec['flops_per_step'] / d['tflops_step'] / 1e9:.3f} "
                 f"ms -> {d['tflops_step']:.1f} TFLOP/s")
    if d["clock_ghz"] is not None:
        lines.append(f"* clock during the kernel: {d['clock_ghz']:.2f} GHz (GRBM_GUI_ACTIVE is "
                     f"summed over the {XCDS} XCDs; clock = GRBM/{XCDS}/duration)")
    if d["mfma_busy"] is not None:
        lines.append(f"* MFMA busy: {d['mfma_busy'] * 100:.1f}% (BUSY / ({SIMDS} SIMDs x cycles "
                     "per XCD))")
    if t:
        lines.append(f"* HBM traffic per step (all launches): {t['hbm_bytes'] / 1e9:.2f} GB "
                     f"(FETCH x {t['scale_fetch']:.3f} + WRITE x {t['scale_write']:.3f}, copy "
                     "calibration)")
    b = rec["bench"]
    if b.get("kernel_ms_avg") is not None:
        lines.append(f"* bench.py line of the same workload: {b['value']} {b['unit']}, launch "
                     f"{b['kernel_ms_avg']} ms (HIP events), frac {b['frac']}")
    return "\n".join(lines) + "\n"


def update_traffic(path, key, workload, rec, command=""):
    """Writes the calibrated HBM bytes of record `rec` into the traffic table
    at `path` (profiles/traffic.json, which bench.py reads for its roofline
    line) under `key` (the bench variant), for the workload string bench.py
    matches (e.g. "N=8192 NITER=10"), together with the copy calibration the
    bytes were scaled by.  Other entries are kept."""
    t = rec.get("traffic")
    if not t:
        raise ValueError("the record has no traffic (FETCH_SIZE / WRITE_SIZE runs and the copy "
                         "calibration are needed)")
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        table = {}
    table["calibration"] = {
        "kernel": COPY_KERNEL + " (1 GiB device-to-device copy)",
        "fetch_kib": (1 << 30) / 1024.0 / t["scale_fetch"],
        "write_kib": (1 << 30) / 1024.0 / t["scale_write"],
        "bytes_each_way": 1 << 30,
        "scale_fetch": round(t["scale_fetch"], 4),
        "scale_write": round(t["scale_write"], 4),
    }
    table[key] = {
        "workload": workload,
        "kernels": rec["step_kernels"],
        "fetch_kib_raw": t["fetch_kib"],
        "write_kib_raw": t["write_kib"],
        "hbm_bytes_per_launch": int(round(t["hbm_bytes"])),
        "command": command,
    }
    with open(path, "w") as f:
        json.dump(table, f, indent=1)
        f.write("\n")
    return table[key]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--bench", required=True, help="file holding the bench.py JSON line")
    ap.add_argument("--stats", required=True, help="rocprofv3 --kernel-trace --stats output")
    ap.add_argument("--pmc", nargs="*", default=[], help="rocprofv3 --pmc outputs")
    ap.add_argument("--copy-fetch", help="--pmc FETCH_SIZE output of bench.py --mode copy")
    ap.add_argument("--copy-write", help="--pmc WRITE_SIZE output of bench.py --mode copy")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--dominant", help="substring of the dominant kernel's name")
    ap.add_argument("--warmup", type=int, default=0,
                    help="leave out each kernel's first N dispatches (the bench's warm-up "
                         "launches), using the --stats run's kernel_trace.csv")
    ap.add_argument("--json", help="write the record here")
    ap.add_argument("--markdown", help="write the Markdown summary here")
    ap.add_argument("--title", default="profiling record")
    ap.add_argument("--command", action="append", default=[],
                    help="command that produced an input (repeatable)")
    ap.add_argument("--traffic-json", help="traffic table to update (profiles/traffic.json)")
    ap.add_argument("--traffic-key", help="its entry: the bench variant")
    ap.add_argument("--traffic-workload", help="the workload string bench.py matches")
    a = ap.parse_args(argv)
    with open(a.bench) as f:
        bench = json.loads([ln for ln in f if ln.strip().startswith("{")][-1])
    scales = None
    if a.copy_fetch and a.copy_write:
        scales = copy_scales(read_counters(a.copy_fetch), read_counters(a.copy_write),
                             a.copy_bytes)
    stats = read_kernel_stats(a.stats)
    if a.warmup > 0:
        stats = skip_warmup(stats, read_kernel_trace(a.stats), a.warmup)
    rec = make_record(bench, stats, read_counters(*a.pmc), a.dominant, scales, a.command)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(rec, f, indent=1)
    if a.markdown:
        with open(a.markdown, "w") as f:
            f.write(to_markdown(rec, a.title))
    if a.traffic_json:
        if not (a.traffic_key and a.traffic_workload):
            ap.error("--traffic-json needs --traffic-key and --traffic-workload")
        cmd = next((c for c in a.command if "FETCH_SIZE" in c), "")
        update_traffic(a.traffic_json, a.traffic_key, a.traffic_workload, rec,
                       cmd + "   (and --pmc WRITE_SIZE)" if cmd else "")
    if not a.json and not a.markdown and not a.traffic_json:
        json.dump(rec, sys.stdout, indent=1)
        print()
    return 0


if __name__ == "__main__":
    sys.exit(main())
