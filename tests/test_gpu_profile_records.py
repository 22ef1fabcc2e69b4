"""The profiling-record parser (hb_hammerbench_amd/profile_records.py, the
counterpart of the reference's vanilla_stats region of interest,
apps/sgemm/kernel.cpp:288,329, and its per-app parser, apps/bfs/parse.py:12-27)
on real rocprofv3 output: bench.py runs under `rocprofv3 --kernel-trace
--stats` and under one `--pmc` pass as a child process, and the record built
from the CSVs must name the GEMM as the dominant kernel, agree with the bench
line's own HIP-event timing, and count exactly the algorithmic flops on the
matrix cores (N = 4096, NITER = 2: the tiles divide N, so no padding)."""
import json
import os
import shutil
import subprocess
import sys

import pytest

from hb_hammerbench_amd import profile_records as pr

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N, NITER = 4096, 2
BENCH = [sys.executable, os.path.join(ROOT, "bench.py"), "--n", str(N), "--niter", str(NITER),
         "--steps", "3", "--warmup", "1", "--no-cpu", "--no-compare", "--no-verify",
         "--no-vendor"]


def rocprof():
    path = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(path):
        pytest.skip("rocprofv3 not installed")
    return path


def run_profiled(args, outdir, timeout):
    """bench.py under rocprofv3, run from /tmp with TMPDIR=/tmp (the profiler
    keeps temporary files there; it must name an existing directory), CSV
    output under `outdir`."""
    env = dict(os.environ, TMPDIR="/tmp")
    cmd = [rocprof()] + args + ["-f", "csv", "-d", str(outdir), "-o", "run", "--"] + BENCH
    print("\n" + " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True,
                       timeout=timeout)
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    # rocprofv3 itself may end with a segmentation fault after a complete run
    assert lines, f"no bench line (rc {r.returncode}): {r.stdout[-1500:]} {r.stderr[-1500:]}"
    return json.loads(lines[-1])


def test_record_from_real_rocprofv3_output(tmp_path):
    bench = run_profiled(["--kernel-trace", "--stats"], tmp_path / "stats", timeout=240)
    stats = pr.read_kernel_stats(str(tmp_path / "stats"))
    rec = pr.make_record(bench, stats, {})
    assert "sgemm_f32_kernel" in rec["dominant"]["name"], rec["step_kernels"]
    assert rec["schema"] == pr.SCHEMA
    dom = [r for r in rec["kernels"] if r["name"] == rec["dominant"]["name"]][0]
    assert dom["calls"] == 3 + 1  # timed steps and the warm-up, one launch each
    d = rec["derived"]
    assert d["tflops_dominant"] == pytest.approx(bench["roofline"]["achieved"], rel=0.05)
    assert dom["avg_ns"] / 1e6 == pytest.approx(bench["roofline"]["kernel_ms_avg"], rel=0.05)
    md = pr.to_markdown(rec, "sgemm N=4096 NITER=2")
    assert rec["dominant"]["name"] in md and "TFLOP/s" in md

    bench2 = run_profiled(["--pmc", "SQ_INSTS_VALU_MFMA_MOPS_F32", "GRBM_GUI_ACTIVE"],
                          tmp_path / "pmc", timeout=240)
    counters = pr.read_counters(str(tmp_path / "pmc"))
    rec2 = pr.make_record(bench2, stats, counters)
    flops = 2.0 * N ** 3 * NITER
    assert rec2["flops_per_step"] == flops
    assert rec2["derived"]["matrix_ops_executed"] == flops
    assert 1.0 < rec2["derived"]["clock_ghz"] < 2.6
