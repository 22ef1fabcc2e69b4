"""Host sanitizer build (CPU only, no GPU, no device code instrumented).

The oracle (oracle/oracle.c) and the grouped planner's host half
(hb_hammerbench_amd/csrc/grouped_plan.cpp, everything hbs_grouped_plan_create
does before it touches the device, and all of hbs_grouped_plan_describe) are
built with -fsanitize=address,undefined by `make asan` in their folders, and
the oracle's tests and the planner's dry-run tests run against those builds
in a child Python with the ASan runtime preloaded.  The reference relies on
fences and barriers alone on this path (apps/sgemm/kernel.cpp:282-283,
327-331); this checks the host code that stands in for its runtime."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "hb_hammerbench_amd", "csrc")
PLANNER = os.path.join(ROOT, "hb_hammerbench_amd", "_asan", "libhbsplan_asan.so")
ORACLE = os.path.join(ROOT, "oracle", "_asan", "liboracle.so")


def asan_runtime():
    try:
        path = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                              text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        path = ""
    if not os.path.isabs(path) or not os.path.exists(path):
        pytest.skip("gcc's ASan runtime is not installed")
    return path


@pytest.fixture(scope="module")
def sanitized_env():
    asan = asan_runtime()
    subprocess.check_call(["make", "-s", "-C", CSRC, "asan"])
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"])
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": asan,
        # Python's own allocations go through malloc, so they carry redzones too
        "PYTHONMALLOC": "malloc",
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:exitcode=23",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "HBS_ORACLE_LIB": ORACLE,
        "HBS_PLANNER_LIB": PLANNER,
    })
    return env


def run(env, args, timeout=900):
    return subprocess.run([sys.executable] + args, cwd=ROOT, env=env, capture_output=True,
                          text=True, timeout=timeout)


def test_oracle_and_planner_tests_pass_under_asan_and_ubsan(sanitized_env):
    r = run(sanitized_env, ["-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                            "tests/test_oracle.py", "tests/test_plan_describe.py"])
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert " passed" in r.stdout


def test_sanitized_builds_are_the_ones_loaded_and_instrumented(sanitized_env):
    """The child process maps the sanitized libraries (not the plain ones),
    and a read past the end of the problem array -- a planted caller mistake,
    count 2 for a malloc'd array of one -- is reported by ASan, so the
    instrumentation is live."""
    probe = (
        "import ctypes, hb_hammerbench_amd as hb, oracle\n"
        "oracle.lib(); hb.describe_plan([(64, 64, 64)])\n"
        "maps = open('/proc/self/maps').read()\n"
        f"print('PLANNER_MAPPED', {PLANNER!r} in maps, 'ORACLE_MAPPED', {ORACLE!r} in maps, "
        "flush=True)\n"
        "libc = ctypes.CDLL(None)\n"
        "libc.malloc.restype = ctypes.c_void_p; libc.malloc.argtypes = [ctypes.c_size_t]\n"
        "one = hb.Problem(16, 16, 16, 8, 8, 8, 8, 8, 8)\n"
        "buf = libc.malloc(ctypes.sizeof(one))\n"
        "ctypes.memmove(buf, ctypes.addressof(one), ctypes.sizeof(one))\n"
        "arr = ctypes.cast(buf, ctypes.POINTER(hb.Problem))\n"
        "info = hb.GroupedInfo()\n"
        "hb.planner().hbs_grouped_plan_describe(arr, 2, 0, None, -1, 1, 256, ctypes.byref(info))\n"
        "print('NOT_CAUGHT', flush=True)\n")
    env = dict(sanitized_env)
    # a wide redzone after every allocation, so that the planted read of the
    # second problem's fields lands in it rather than in a neighbouring block
    env["ASAN_OPTIONS"] += ":redzone=256"
    r = run(env, ["-c", probe], timeout=300)
    assert "PLANNER_MAPPED True ORACLE_MAPPED True" in r.stdout, r.stdout + r.stderr[-2000:]
    assert r.returncode == 23 and "heap-buffer-overflow" in r.stderr, r.stderr[-3000:]
    assert "NOT_CAUGHT" not in r.stdout
