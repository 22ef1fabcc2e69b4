"""Profiling records (hb_hammerbench_amd/profile_records.py) on synthetic
rocprofv3 CSVs whose derived figures have closed forms: a dominant kernel of
8 ms per dispatch at a 2.4 GHz clock and 90% MFMA busy, byte counters scaled by
a copy calibration of 2 (FETCH_SIZE) and 1 (WRITE_SIZE)."""
import csv
import json
import os

import pytest

from hb_hammerbench_amd import profile_records as pr

KERNEL = "void (anonymous namespace)::sgemm_f32_kernel<256, 256, 32, 4, 2, 0>((anonymous namespace)::GemmArgs)"
FILL = "fill_random_kernel(float*, long, int, unsigned long, unsigned long)"
DOM_NS = 8.0e6
GRBM = pr.XCDS * 2.4e9 * DOM_NS * 1e-9


def write_stats(path, rows):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "run_kernel_stats.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs",
                    "StdDev"])
        for name, calls, avg in rows:
            w.writerow([name, calls, calls * avg, avg, 50.0, avg * 0.9, avg * 1.1, 0.0])


def write_counters(path, rows):
    """rows: (kernel, counter, [value per dispatch])"""
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        d = 0
        for kernel, counter, values in rows:
            for v in values:
                d += 1
                w.writerow([d, kernel, counter, v])


@pytest.fixture
def run_dir(tmp_path):
    t = str(tmp_path)
    write_stats(os.path.join(t, "stats"), [(KERNEL, 6, DOM_NS), (FILL, 2, 1.0e6)])
    busy = 0.9 * pr.SIMDS * GRBM / pr.XCDS
    write_counters(os.path.join(t, "fetch"), [(KERNEL, "FETCH_SIZE", [900.0, 1100.0]),
                                              (FILL, "FETCH_SIZE", [5.0])])
    write_counters(os.path.join(t, "write"), [(KERNEL, "WRITE_SIZE", [500.0, 500.0])])
    write_counters(os.path.join(t, "sq"), [(KERNEL, "GRBM_GUI_ACTIVE", [GRBM, GRBM]),
                                           (KERNEL, "SQ_VALU_MFMA_BUSY_CYCLES", [busy, busy]),
                                           (KERNEL, "SQ_INSTS_VALU_MFMA_MOPS_F32", [1e9, 1e9])])
    write_counters(os.path.join(t, "cfetch"), [("__amd_rocclr_copyBuffer", "FETCH_SIZE", [524288.0])])
    write_counters(os.path.join(t, "cwrite"), [("__amd_rocclr_copyBuffer", "WRITE_SIZE", [1048576.0])])
    bench = {"value": 140.0, "unit": "TFLOP/s",
             "config": {"workload": "synthetic", "variant": "fp32", "flops_per_step": 1.0e12,
                        "bytes_algorithmic": 8.0e9},
             "roofline": {"peak": 157.3, "kernel_ms_avg": 7.9, "frac": 0.9}}
    with open(os.path.join(t, "bench.json"), "w") as f:
        f.write("some log line\n" + json.dumps(bench) + "\n")
    return t, bench


def test_short_name_strips_arguments_and_namespace():
    assert pr.short_name(KERNEL) == "sgemm_f32_kernel<256, 256, 32, 4, 2, 0>"
    assert pr.short_name(FILL) == "fill_random_kernel"


def test_record_derived_figures(run_dir):
    t, bench = run_dir
    stats = pr.read_kernel_stats(os.path.join(t, "stats"))
    ctr = pr.read_counters(*(os.path.join(t, d) for d in ("fetch", "write", "sq")))
    scales = pr.copy_scales(pr.read_counters(os.path.join(t, "cfetch")),
                            pr.read_counters(os.path.join(t, "cwrite")))
    assert scales["scale_fetch"] == 2.0 and scales["scale_write"] == 1.0
    rec = pr.make_record(bench, stats, ctr, None, scales, ["cmd"])
    assert rec["schema"] == pr.SCHEMA
    assert rec["step_kernels"] == ["sgemm_f32_kernel<256, 256, 32, 4, 2, 0>"]
    d = rec["derived"]
    assert d["tflops_dominant"] == pytest.approx(125.0)
    assert d["frac_dominant"] == pytest.approx(125.0 / 157.3)
    assert d["clock_ghz"] == pytest.approx(2.4)
    assert d["mfma_busy"] == pytest.approx(0.9)
    assert d["matrix_ops_executed"] == pytest.approx(1e9 * 512)
    # per-dispatch means, step kernels only (the fill kernel's bytes are not in a step)
    assert rec["traffic"]["fetch_kib"] == pytest.approx(1000.0)
    assert rec["traffic"]["hbm_bytes"] == pytest.approx((1000.0 * 2 + 500.0 * 1) * 1024)
    assert rec["traffic"]["bytes_algorithmic"] == 8.0e9


def test_dominant_by_name_and_errors(run_dir):
    t, bench = run_dir
    stats = pr.read_kernel_stats(os.path.join(t, "stats"))
    rec = pr.make_record(bench, stats, {}, "sgemm_f32")
    assert rec["dominant"]["avg_ns"] == DOM_NS
    assert rec["derived"]["clock_ghz"] is None and rec["traffic"] is None
    with pytest.raises(ValueError):
        pr.make_record(bench, stats, {}, "no_such_kernel")
    with pytest.raises(FileNotFoundError):
        pr.read_kernel_stats(os.path.join(t, "fetch"))


def test_cli_writes_json_and_markdown(run_dir):
    t, _ = run_dir
    out_json, out_md = os.path.join(t, "rec.json"), os.path.join(t, "rec.md")
    rc = pr.main(["--bench", os.path.join(t, "bench.json"), "--stats", os.path.join(t, "stats"),
                  "--pmc", os.path.join(t, "fetch"), os.path.join(t, "write"), os.path.join(t, "sq"),
                  "--copy-fetch", os.path.join(t, "cfetch"), "--copy-write", os.path.join(t, "cwrite"),
                  "--dominant", "sgemm_f32_kernel", "--json", out_json, "--markdown", out_md,
                  "--title", "synthetic", "--command", "rocprofv3 ..."])
    assert rc == 0
    rec = json.load(open(out_json))
    assert rec["schema"] == pr.SCHEMA and rec["commands"] == ["rocprofv3 ..."]
    md = open(out_md).read()
    assert md.startswith("# synthetic\n")
    assert "**125.0 TFLOP/s** = 79.5% of 157.3" in md
    assert "MFMA busy: 90.0%" in md and "clock during the kernel: 2.40 GHz" in md
    assert "HBM traffic per step (all launches): 0.00 GB (FETCH x 2.000 + WRITE x 1.000" in md


def test_warmup_dispatches_left_out_of_the_averages(tmp_path):
    """kernel_trace.csv durations, in start order: the first `warmup`
    dispatches of each kernel (the bench's untimed warm-up) are dropped from
    calls, average, minimum and maximum; kernels launched fewer times keep
    their kernel_stats.csv figures."""
    d = str(tmp_path / "stats")
    write_stats(d, [(KERNEL, 5, 7.0e6), (FILL, 2, 1.0e6)])
    durs = [12.0e6, 9.0e6, 6.0e6, 5.0e6, 4.0e6]  # kernel dispatches in start order
    with open(os.path.join(d, "run_kernel_trace.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kind", "Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        t = 1000
        rows = []
        for dur in durs:
            rows.append(["KERNEL_DISPATCH", KERNEL, t, t + int(dur)])
            t += int(dur) + 100
        rows.insert(1, ["KERNEL_DISPATCH", FILL, 5, 5 + 1000000])
        rows.append(["KERNEL_DISPATCH", FILL, t, t + 1000000])
        for r in reversed(rows):  # file order must not matter
            w.writerow(r)
    trace = pr.read_kernel_trace(d)
    name = pr.short_name(KERNEL)
    assert trace[name] == [int(x) for x in durs]
    stats = pr.skip_warmup(pr.read_kernel_stats(d), trace, 2)
    dom = [r for r in stats if r["name"] == name][0]
    assert dom["calls"] == 3 and dom["avg_ns"] == pytest.approx(5.0e6)
    assert dom["min_ns"] == 4.0e6 and dom["max_ns"] == 6.0e6
    fill = [r for r in stats if r["name"] == pr.short_name(FILL)][0]
    assert fill["calls"] == 2 and fill["avg_ns"] == 1.0e6


def test_traffic_table_update_keeps_other_entries(run_dir, tmp_path):
    """--traffic-json writes the calibrated bytes of the record under the
    bench variant and the workload string bench.py looks up, with the copy
    calibration, and keeps the other entries."""
    t, _ = run_dir
    table = str(tmp_path / "traffic.json")
    with open(table, "w") as f:
        json.dump({"fp8res": {"workload": "M=1 N=1 K=1", "hbm_bytes_per_launch": 7}}, f)
    rc = pr.main(["--bench", os.path.join(t, "bench.json"), "--stats", os.path.join(t, "stats"),
                  "--pmc", os.path.join(t, "fetch"), os.path.join(t, "write"),
                  "--copy-fetch", os.path.join(t, "cfetch"), "--copy-write", os.path.join(t, "cwrite"),
                  "--traffic-json", table, "--traffic-key", "fp32",
                  "--traffic-workload", "N=8192 NITER=10",
                  "--command", "rocprofv3 --pmc FETCH_SIZE ...", "--command", "rocprofv3 --pmc WRITE_SIZE ..."])
    assert rc == 0
    got = json.load(open(table))
    assert got["fp8res"]["hbm_bytes_per_launch"] == 7
    e = got["fp32"]
    assert e["workload"] == "N=8192 NITER=10" and e["kernels"] == ["sgemm_f32_kernel<256, 256, 32, 4, 2, 0>"]
    assert e["hbm_bytes_per_launch"] == (1000 * 2 + 500) * 1024
    assert e["command"].startswith("rocprofv3 --pmc FETCH_SIZE")
    assert got["calibration"]["scale_fetch"] == 2.0 and got["calibration"]["scale_write"] == 1.0
    with pytest.raises(SystemExit):
        pr.main(["--bench", os.path.join(t, "bench.json"), "--stats", os.path.join(t, "stats"),
                 "--traffic-json", table])
