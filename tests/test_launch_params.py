"""Launch-parameter generation and launcher argument handling (no GPU).

Counterpart of apps/sgemm/Makefile:7-28 + mk/testbench_common.mk:12-29: the
test name N_512__NITER_2 (apps/sgemm/tests.mk:1) yields parameters.mk with
N = 512 and NITER = 2; optional fields select precision and input kind.
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
                 "PREC": "fp32", "INPUT": "pattern"}
    q = _params(tmp_path / "N_8192__NITER_1__PREC_bf16x3__INPUT_loguniform" / "parameters.mk")
    assert (q["N"], q["NITER"], q["PREC"], q["INPUT"]) == ("8192", "1", "bf16x3", "loguniform")
    assert (tmp_path / "N_512__NITER_2" / "Makefile").read_text() == \
        open(os.path.join(APP, "template.mk")).read()
    subprocess.check_call(["make", "-s", "-C", str(tmp_path), "purge"])
    assert not (tmp_path / "N_512__NITER_2").exists()


LAUNCHER = os.path.join(APP, "sgemm")


@pytest.mark.skipif(not os.path.exists(LAUNCHER), reason="launcher not built")
@pytest.mark.parametrize("args", [["--N", "0"], ["--NITER", "-1"], ["--precision", "fp16"],
                                  ["--input", "zeros"], ["--golden", "x"], ["--bogus"]])
def test_launcher_rejects_bad_arguments(args):
    r = subprocess.run([LAUNCHER] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
