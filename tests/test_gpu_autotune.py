"""The grouped autotuner (hb_hammerbench_amd/apps/autotune) end to end on the
GPU.  The reference fixes its tile-group shape at build time
(apps/sgemm/template.mk:7-9) and sweeps shapes by hand
(apps/legacy/gemm/sgemm_512/tests.mk:1-6); here `autotune --group` times every
tile configuration (shape, wave count and arrangement, K stage, launch) and
split-K per size class of a committed 16-problem list, and the tuning file it
writes must be accepted by hbs_grouped_plan_create and give results within
the FP64 bounds."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle

torch = pytest.importorskip("torch")
import hb_hammerbench_amd as hb  # noqa: E402

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TUNER = os.path.join(ROOT, "hb_hammerbench_amd", "apps", "autotune", "autotune")
SHAPES = os.path.join(ROOT, "tests", "golden", "autotune_16_shapes.txt")
U = 2.0 ** -24


def gamma(n):
    return n * U / (1 - n * U)


def test_autotune_group_writes_a_usable_tuning_file(tmp_path):
    assert os.path.exists(TUNER), "autotuner not built"
    out = tmp_path / "tune.txt"
    r = subprocess.run([TUNER, "--group", SHAPES, "--reps", "2", "--out", str(out)],
                       capture_output=True, text=True, timeout=900)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    shapes = [tuple(int(v) for v in ln.split()) for ln in open(SHAPES) if ln.strip()]
    text = out.read_text()
    # one margin line per visited class and precision: the fastest candidate
    # and the time the kept table gives
    margins = re.findall(r"^#   (fp32|bf16x3) class (\d+) (\d+) (\d+) .* fastest candidate tile "
                         r"(\d+)", text, re.M)
    assert {m[0] for m in margins} == {"fp32", "bf16x3"}
    assert len(margins) >= 8
    entries = [ln.split() for ln in text.splitlines() if ln.strip() and not ln.startswith("#")]
    for e in entries:
        prec, mc, nc, kc, tile, split = e[0], *map(int, e[1:6])
        assert prec in ("fp32", "bf16x3") and 0 <= tile <= 9 and 1 <= split <= 4
        assert all(c in (64, 128, 256, 512, 1024, 2048) for c in (mc, nc, kc))
    # the planner accepts the file; both precisions within the FP64 bounds
    ta = sum(M * K for M, N, K in shapes); tb = sum(K * N for M, N, K in shapes)
    tc = sum(M * N for M, N, K in shapes)
    a = torch.empty(ta, device="cuda"); b = torch.empty(tb, device="cuda")
    c = torch.empty(tc, device="cuda")
    hb.fill_random(a, ta, hb.GEN_UNIFORM, 3); hb.fill_random(b, tb, hb.GEN_UNIFORM, 4)
    probs, views, oa, ob, oc = [], [], 0, 0, 0
    for M, N, K in shapes:
        A = a[oa:oa + M * K].view(M, K); B = b[ob:ob + K * N].view(K, N)
        C = c[oc:oc + M * N].view(M, N)
        probs.append((A, B, C, M, N, K)); views.append((A, B, C, M, N, K))
        oa += M * K; ob += K * N; oc += M * N
    for prec in (hb.PREC_FP32, hb.PREC_BF16X3):
        c.fill_(float("nan"))
        plan = hb.GroupedPlan(probs, prec, tuning_file=str(out))
        g = plan.graph()
        g.launch()
        torch.cuda.synchronize()
        for A, B, C, M, N, K in views:
            R, S = oracle.gemm_f64(A.cpu().numpy(), B.cpu().numpy())
            bound = (gamma(K + 2) if prec == hb.PREC_FP32
                     else 3.1 * 2.0 ** -16 + gamma(3 * K + 2)) * S
            assert np.all(np.abs(C.cpu().numpy() - R) <= bound), (prec, M, N, K)
        g.close()
        plan.close()
