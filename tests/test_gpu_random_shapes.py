"""Seeded random shapes, leading dimensions and alignments through every GEMM
entry point of the C-ABI, against the FP64 oracle.

The fixed-shape tests pin each tile mode on chosen cases; these draw the
cases instead: M, N, K from 1 to a few hundred (mostly not multiples of any
tile or stage), leading dimensions padded by 0..7 elements, and operand
pointers offset by 0..3 floats, so that the interior, clamped, K-predicated
and per-element predicated modes, edge tiles of every size, and the grouped
launches are all reached.  Bars, per element:
 * FP32 (hbs_sgemm_strided): |C - R| <= gamma_{K+2} S;
 * BF16x3 (hbs_gemm_bf16x3): |C - R| <= (3.1 * 2^-16 + gamma_{3K+2}) S;
 * grouped plans of both precisions: the same per problem;
 * stream-K (hbs_gemm_streamk) of both precisions: the same with s more
   additions for a tile split s ways (s <= G + 1, G its workgroups), through
   one workspace reused by every case;
with R and S = |A||B| from the oracle's FP64 product, and every element of C
outside the M x N result (padding columns and the gaps around it) unchanged.
"""
import numpy as np
import pytest

import oracle

torch = pytest.importorskip("torch")
import hb_hammerbench_amd as hb  # noqa: E402

pytestmark = pytest.mark.gpu
U = 2.0 ** -24
SENTINEL = -5.0


def gamma(n):
    return n * U / (1 - n * U)


def bound(prec, K, S, extra=0):
    if prec == hb.PREC_FP32:
        return gamma(K + 2 + extra) * S
    return (3.1 * 2.0 ** -16 + gamma(3 * K + 2 + extra)) * S


def draw_case(rng, big=False):
    """(M, N, K, pads, offsets): sizes from 1 to 320 (to 700 for `big`), most
    of them ragged."""
    hi = 700 if big else 320
    M, N, K = (int(x) for x in rng.integers(1, hi + 1, 3))
    pads = tuple(int(x) for x in rng.integers(0, 8, 3))
    offs = tuple(int(x) for x in rng.integers(0, 4, 3))
    return M, N, K, pads, offs


def operands(case, seed):
    """Device buffers holding A (M x K, row stride lda), B (K x N, ldb) and a
    sentinel-filled C (M x N, ldc), each starting `off` floats into its buffer."""
    M, N, K, (pa, pb, pc), (oa, ob, oc) = case
    lda, ldb, ldc = K + pa, N + pb, N + pc
    a = torch.empty(oa + M * lda, dtype=torch.float32, device="cuda")
    b = torch.empty(ob + K * ldb, dtype=torch.float32, device="cuda")
    c = torch.full((oc + M * ldc + 5,), SENTINEL, dtype=torch.float32, device="cuda")
    hb.fill_random(a, a.numel(), hb.GEN_UNIFORM, seed)
    hb.fill_random(b, b.numel(), hb.GEN_LOGUNIFORM, seed + 1)
    return a, b, c, lda, ldb, ldc


def check(prec, case, a, b, c, lda, ldb, ldc, extra=0):
    M, N, K, _, (oa, ob, oc) = case
    Ah, Bh, Ch = a.cpu().numpy(), b.cpu().numpy(), c.cpu().numpy()
    A = np.lib.stride_tricks.as_strided(Ah[oa:], (M, K), (lda * 4, 4)).copy()
    B = np.lib.stride_tricks.as_strided(Bh[ob:], (K, N), (ldb * 4, 4)).copy()
    C = np.lib.stride_tricks.as_strided(Ch[oc:], (M, N), (ldc * 4, 4)).copy()
    R, S = oracle.gemm_f64(A, B)
    err = np.abs(C.astype(np.float64) - R)
    assert np.all(err <= bound(prec, K, S, extra)), (case, float(err.max()))
    written = np.zeros(Ch.size, bool)
    for r in range(M):
        written[oc + r * ldc: oc + r * ldc + N] = True
    assert np.all(Ch[~written] == SENTINEL), case


def cases(seed, n, big_every=4):
    rng = np.random.default_rng(seed)
    return [draw_case(rng, big=(i % big_every == 0)) for i in range(n)]


@pytest.mark.parametrize("i,case", list(enumerate(cases(2026, 32))))
def test_random_shapes_fp32_strided(i, case):
    print(f"fp32 case {i}: {case}", flush=True)
    a, b, c, lda, ldb, ldc = operands(case, 100 + i)
    M, N, K, _, (oa, ob, oc) = case
    hb.sgemm_strided(a[oa:], b[ob:], c[oc:], M, N, K, lda, ldb, ldc)
    torch.cuda.synchronize()
    check(hb.PREC_FP32, case, a, b, c, lda, ldb, ldc)


@pytest.mark.parametrize("i,case", list(enumerate(cases(3027, 32))))
def test_random_shapes_bf16x3(i, case):
    print(f"bf16x3 case {i}: {case}", flush=True)
    a, b, c, lda, ldb, ldc = operands(case, 200 + i)
    M, N, K, _, (oa, ob, oc) = case
    hb.gemm_bf16x3(a[oa:], b[ob:], c[oc:], M, N, K, lda=lda, ldb=ldb, ldc=ldc)
    torch.cuda.synchronize()
    check(hb.PREC_BF16X3, case, a, b, c, lda, ldb, ldc)


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
@pytest.mark.parametrize("seed", [11, 12, 13])
def test_random_grouped_plans(prec, seed):
    """One plan of 24 random problems: some with 16-byte-loadable operands
    (clamped launches), some not (the predicated launch), K both multiples of
    32 and not; replayed twice from a hipGraph, both replays bit-identical."""
    rng = np.random.default_rng(seed)
    probs_cases = [draw_case(rng, big=(j % 6 == 0)) for j in range(24)]
    print(f"grouped prec {prec} seed {seed}: {probs_cases}", flush=True)
    bufs, probs = [], []
    for j, case in enumerate(probs_cases):
        a, b, c, lda, ldb, ldc = operands(case, 300 + 10 * seed + j)
        M, N, K, _, (oa, ob, oc) = case
        bufs.append((case, a, b, c, lda, ldb, ldc))
        probs.append((a[oa:], b[ob:], c[oc:], M, N, K, lda, ldb, ldc))
    plan = hb.GroupedPlan(probs, prec)
    graph = plan.graph()
    graph.launch()
    torch.cuda.synchronize()
    first = [c.clone() for (_, _, _, c, _, _, _) in bufs]
    graph.launch()
    torch.cuda.synchronize()
    for (case, a, b, c, lda, ldb, ldc), c0 in zip(bufs, first):
        assert torch.equal(c.view(torch.int32), c0.view(torch.int32)), case
        check(prec, case, a, b, c, lda, ldb, ldc)
    graph.close()
    plan.close()


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
def test_random_shapes_streamk(prec):
    """24 random cases through one stream-K workspace: each call must leave
    the tile counters zero for the next, whatever its shape."""
    ws = torch.zeros(hb.streamk_workspace_size(8192, 8192, 8192, prec), dtype=torch.uint8,
                     device="cuda")
    for i, case in enumerate(cases(4100 + prec, 24, big_every=3)):
        print(f"streamk prec {prec} case {i}: {case}", flush=True)
        a, b, c, lda, ldb, ldc = operands(case, 400 + i)
        M, N, K, _, (oa, ob, oc) = case
        hb.gemm_streamk(a[oa:], b[ob:], c[oc:], M, N, K, prec, lda, ldb, ldc, workspace=ws)
        torch.cuda.synchronize()
        G = (hb.streamk_workspace_size(M, N, K, prec) - 256 * 1024) // (2 * 256 * 256 * 4)
        check(prec, case, a, b, c, lda, ldb, ldc, extra=G + 1)
        assert not ws[:256 * 1024].any().item(), case
