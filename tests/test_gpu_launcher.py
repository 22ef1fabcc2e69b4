"""End-to-end runs of the C++ launcher (hb_hammerbench_amd/apps/sgemm/sgemm),
the counterpart of running the reference's N_512__NITER_2 test: it initialises,
computes the host golden, launches, reads back and applies the SSE rule
(apps/sgemm/main.cpp:30-149), ending in a regression verdict line."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "hb_hammerbench_amd", "apps", "sgemm")
LAUNCHER = os.path.join(APP, "sgemm")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("args", [
    ["--N", "512", "--NITER", "2"],                                   # N_512__NITER_2
    ["--N", "512", "--NITER", "2", "--input", "uniform"],
    ["--N", "512", "--NITER", "2", "--precision", "bf16x3"],
    # log-uniform values reach 1e6: the absolute SSE rule only fits the pattern,
    # so these inputs are checked against FP64 (componentwise gate)
    ["--N", "300", "--NITER", "3", "--input", "loguniform", "--precision", "bf16x3",
     "--golden", "f64"],
    ["--N", "2048", "--NITER", "2", "--input", "uniform", "--golden", "f64", "--reps", "3"],
    ["--N", "4096", "--NITER", "1", "--golden", "f64"],               # pattern via FP64 check
    # the slices as one grouped plan (hipGraph): default, forced tile and
    # split-K, the committed tuning file
    ["--N", "512", "--NITER", "2", "--mode", "grouped"],
    ["--N", "640", "--NITER", "3", "--mode", "grouped", "--precision", "bf16x3", "--tile", "0",
     "--split", "2"],
    ["--N", "300", "--NITER", "2", "--mode", "grouped", "--tile", "1", "--split", "3",
     "--input", "uniform", "--golden", "f64", "--reps", "2"],
    # the other tile configurations: 8-wave 256 x 128 large tiles, and the
    # FP32-only 4 x 1-wave 128 x 128 tile with split-K
    ["--N", "640", "--NITER", "2", "--mode", "grouped", "--precision", "bf16x3", "--tile", "8"],
    ["--N", "300", "--NITER", "2", "--mode", "grouped", "--tile", "5", "--split", "2",
     "--input", "uniform", "--golden", "f64"],
    # N = 301: slices start off 16-byte boundaries and N is odd, so the
    # grouped plan runs them on its predicated large-tile launch
    ["--N", "301", "--NITER", "3", "--mode", "grouped", "--input", "uniform", "--golden", "f64"],
    ["--N", "301", "--NITER", "2", "--mode", "grouped", "--precision", "bf16x3"],
    ["--N", "512", "--NITER", "2", "--mode", "grouped", "--tuning",
     os.path.join(ROOT, "hb_hammerbench_amd", "tuning", "mi355x_grouped.txt")],
])
def test_launcher_passes(args):
    assert os.path.exists(LAUNCHER), "launcher not built"
    r = subprocess.run([LAUNCHER] + args, capture_output=True, text=True, timeout=240)
    print(r.stdout[-600:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert r.stdout.strip().splitlines()[-1] == "HBS REGRESSION TEST PASSED"


def test_make_generate_and_exec_reference_test(tmp_path):
    """make generate + make N_512__NITER_2.exec, as in the reference workflow."""
    for f in ["Makefile", "tests.mk", "template.mk"]:
        with open(os.path.join(APP, f)) as src:
            (tmp_path / f).write_text(src.read())
    (tmp_path / "app_path.mk").write_text(f"APP_PATH := {APP}\n")
    subprocess.check_call(["make", "-s", "-C", str(tmp_path), "generate"], stdout=subprocess.DEVNULL)
    r = subprocess.run(["make", "-s", "-C", str(tmp_path), "N_512__NITER_2.exec"],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    log = (tmp_path / "N_512__NITER_2" / "exec.log").read_text()
    assert "N=512" in log and "NITER=2" in log and "HBS REGRESSION TEST PASSED" in log
    name = "N_640__NITER_3__MODE_grouped__PREC_bf16x3__TILE_0__SPLIT_2"
    r = subprocess.run(["make", "-s", "-C", str(tmp_path), name + ".exec"],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    log = (tmp_path / name / "exec.log").read_text()
    assert "mode=grouped" in log and "tile=0 split=2" in log
    assert "HBS REGRESSION TEST PASSED" in log


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_launcher_fails_on_a_planted_golden_mismatch(precision):
    """The validation of main.cpp:120-134 itself: one golden element moved by
    0.5 on the host gives SSE = 0.25 >= 0.01, so the launcher must print the
    mismatch, end in HBS REGRESSION TEST FAILED and exit with status 1; moved by
    0.05 (SSE = 0.0025) the element still differs but the rule passes."""
    base = [LAUNCHER, "--N", "256", "--NITER", "2", "--precision", precision]
    r = subprocess.run(base + ["--perturb-golden", "70000:0.5"], capture_output=True, text=True,
                       timeout=240)
    print(r.stdout[-600:])
    assert r.returncode == 1, r.stdout[-2000:] + r.stderr[-2000:]
    out = r.stdout.strip().splitlines()
    assert out[-1] == "HBS REGRESSION TEST FAILED"
    assert "elements differing: 1, SSE= 0.250000" in r.stdout
    assert "Matrix Mismatch" in r.stdout and "[70000] actual=" in r.stdout
    r = subprocess.run(base + ["--perturb-golden", "70000:0.05"], capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "elements differing: 1," in r.stdout
    assert r.stdout.strip().splitlines()[-1] == "HBS REGRESSION TEST PASSED"


def test_launcher_host_code_under_asan_and_ubsan():
    """The launcher with ASan + UBSan on its host code (`make -C
    hb_hammerbench_amd/apps asan`): the reference configuration N_512__NITER_2
    with the host golden and SSE rule (apps/sgemm/main.cpp:15-26,120-134) runs
    clean, and the planted golden mismatch still fails with status 1 -- the
    argument parsing, golden, readback and validation code raise no report."""
    prog = os.path.join(APP, "sgemm_asan")
    assert os.path.exists(prog), "sanitized launcher not built"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:exitcode=23",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    for extra, code in (([], 0), (["--perturb-golden", "1000:0.5"], 1),
                        (["--precision", "bf16x3", "--mode", "grouped"], 0)):
        r = subprocess.run([prog, "--N", "512", "--NITER", "2"] + extra, capture_output=True,
                           text=True, timeout=240, env=env)
        out = r.stdout + r.stderr
        print(out[-400:])
        assert "AddressSanitizer" not in out and "runtime error" not in out, out[-3000:]
        assert r.returncode == code, out[-2000:]
