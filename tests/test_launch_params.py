"""Launch-parameter generation and launcher argument handling (no GPU).

Counterpart of apps/sgemm/Makefile:7-28 + mk/testbench_common.mk:12-29: the
test name N_512__NITER_2 (apps/sgemm/tests.mk:1) yields parameters.mk with
N = 512 and NITER = 2; optional fields select precision, input kind and the
grouped mode with its tile shape, split-K or tuning file.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "hb_hammerbench_amd", "apps", "sgemm")


def _params(path):
    out = {}
    for line in open(path):
        k, _, v = line.partition("=")
        out[k.strip()] = v.strip()
    return out


def test_make_generate_creates_reference_test_dir(tmp_path):
    for f in ["Makefile", "tests.mk", "template.mk", "app_path.mk"]:
        shutil.copy(os.path.join(APP, f), tmp_path / f)
    subprocess.check_call(["make", "-s", "-C", str(tmp_path), "generate"],
                          stdout=subprocess.DEVNULL)
    p = _params(tmp_path / "N_512__NITER_2" / "parameters.mk")
    assert p == {"test-name": "N_512__NITER_2", "N": "512", "NITER": "2",
                 "PREC": "fp32", "INPUT": "pattern", "MODE": "square", "TILE": "auto",
                 "SPLIT": "1", "TUNING": "none"}
    q = _params(tmp_path / "N_8192__NITER_1__PREC_bf16x3__INPUT_loguniform" / "parameters.mk")
    assert (q["N"], q["NITER"], q["PREC"], q["INPUT"]) == ("8192", "1", "bf16x3", "loguniform")
    g = _params(tmp_path / "N_640__NITER_3__MODE_grouped__PREC_bf16x3__TILE_0__SPLIT_2" /
                "parameters.mk")
    assert (g["N"], g["PREC"], g["MODE"], g["TILE"], g["SPLIT"], g["TUNING"]) == \
        ("640", "bf16x3", "grouped", "0", "2", "none")
    t = _params(tmp_path / "N_512__NITER_2__MODE_grouped__TUNING_mi355x" / "parameters.mk")
    assert (t["MODE"], t["TILE"], t["TUNING"]) == ("grouped", "auto", "mi355x")
    # the per-test Makefile turns the fields into launcher arguments
    args = {}
    for name in ["N_512__NITER_2", "N_640__NITER_3__MODE_grouped__PREC_bf16x3__TILE_0__SPLIT_2",
                 "N_512__NITER_2__MODE_grouped__TUNING_mi355x"]:
        out = subprocess.run(["make", "-s", "-n", "-C", str(tmp_path / name), "exec.log",
                              "SGEMM=/bin/true", f"APP_PATH={APP}"],
                             capture_output=True, text=True, check=True).stdout
        args[name] = out.splitlines()[0].split("2>&1")[0].split()[1:]
    assert "--mode" not in args["N_512__NITER_2"]
    a = args["N_640__NITER_3__MODE_grouped__PREC_bf16x3__TILE_0__SPLIT_2"]
    assert a[a.index("--mode") + 1] == "grouped" and a[a.index("--tile") + 1] == "0"
    assert a[a.index("--split") + 1] == "2" and "--tuning" not in a
    b = args["N_512__NITER_2__MODE_grouped__TUNING_mi355x"]
    tuning = b[b.index("--tuning") + 1]
    assert os.path.exists(tuning) and tuning.endswith("tuning/mi355x_grouped.txt")
    assert (tmp_path / "N_512__NITER_2" / "Makefile").read_text() == \
        open(os.path.join(APP, "template.mk")).read()
    subprocess.check_call(["make", "-s", "-C", str(tmp_path), "purge"])
    assert not (tmp_path / "N_512__NITER_2").exists()


LAUNCHER = os.path.join(APP, "sgemm")


@pytest.mark.skipif(not os.path.exists(LAUNCHER), reason="launcher not built")
@pytest.mark.parametrize("args", [["--N", "0"], ["--NITER", "-1"], ["--precision", "fp16"],
                                  ["--input", "zeros"], ["--golden", "x"], ["--bogus"],
                                  ["--mode", "batched"], ["--tile", "1"],
                                  ["--mode", "grouped", "--tile", "10"],
                                  ["--mode", "grouped", "--precision", "bf16x3", "--tile", "5"],
                                  ["--mode", "grouped", "--split", "0"],
                                  ["--mode", "grouped", "--tile", "1", "--tuning", "t.txt"]])
def test_launcher_rejects_bad_arguments(args):
    r = subprocess.run([LAUNCHER] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_launcher_rejects_bad_golden_perturbation():
    """--perturb-golden needs index:delta, the host golden and an index inside
    the NITER slices; anything else is a usage error (status 2) before any
    device call."""
    import subprocess
    launcher = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "hb_hammerbench_amd", "apps", "sgemm", "sgemm")
    if not os.path.exists(launcher):
        pytest.skip("launcher not built")
    for extra in (["--perturb-golden", "5"], ["--golden", "f64", "--perturb-golden", "5:1"],
                  ["--N", "4", "--NITER", "1", "--perturb-golden", "16:1"],
                  ["--perturb-golden", "-1:1"]):
        r = subprocess.run([launcher] + extra, capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, extra
        assert "usage:" in r.stderr
