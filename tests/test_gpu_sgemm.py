"""GPU parity of the plain FP32-MFMA SGEMM against the oracle, through the C-ABI.

Bars: bit-exact on the reference's integer pattern (apps/sgemm/main.cpp:46-52,
every partial sum is an integer < 2^24); on random inputs every element within
the deterministic FP32 summation bound |C - R| <= gamma_{K+2} (|A||B|) of the
FP64 oracle R (gamma_n = n u / (1 - n u), u = 2^-24), which any order of FP32
additions satisfies, plus the reference's own SSE < 0.01 check against the
host_mm golden (main.cpp:120-134).
"""
import numpy as np
import pytest

import oracle

torch = pytest.importorskip("torch")
import hb_hammerbench_amd as hb  # noqa: E402

pytestmark = pytest.mark.gpu
U = 2.0 ** -24


def gamma(n):
    return n * U / (1 - n * U)


def dev_empty(n):
    return torch.empty(int(n), dtype=torch.float32, device="cuda")


def dev_pattern(N, niter):
    a = dev_empty(niter * N * N); b = dev_empty(niter * N * N)
    hb.fill_pattern(a, b, niter * N * N)
    return a, b


def dev_uniform(n, seed, first=0):
    x = dev_empty(n)
    hb.fill_random(x, n, hb.GEN_UNIFORM, seed, first)
    return x


def pattern_closed_form_gpu(N, it):
    """Closed form of the main.cpp:46-52 product as a rank-21 product of small
    integer matrices (exact in float64): C = P . Q with P[y][r] = cnt(r) *
    ((o + N y + r) mod 7) and Q[r][x] = ((o + x) + (N mod 3) r) mod 3."""
    o = it * N * N
    r = torch.arange(21, dtype=torch.int64, device="cuda")
    cnt = N // 21 + (r < N % 21).to(torch.int64)
    y = torch.arange(N, dtype=torch.int64, device="cuda")
    P = cnt[None, :] * ((o + N * y[:, None] + r[None, :]) % 7)
    Q = ((o + y[None, :]) + (N % 3) * r[:, None]) % 3
    return (P.double() @ Q.double()).float()


def special_inputs(M=96, N=96, K=96, seed=41):
    """Uniform inputs with planted NaN and infinities: lone +-Inf and NaN, an
    Inf times an exact zero (NaN), and +Inf and -Inf products in one sum (NaN);
    R is the product with IEEE semantics in float64 (elementwise, so no library
    shortcut can skip a zero)."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    A[3, 5] = np.inf; A[10, 7] = -np.inf; A[20, 1] = np.nan
    B[2, 30] = np.inf
    A[60, 40] = np.inf; B[40, 70] = 0.0                        # Inf * 0
    A[80, 11] = np.inf; A[80, 12] = np.inf; B[11, 90] = 1.0; B[12, 90] = -1.0  # Inf - Inf
    with np.errstate(invalid="ignore", over="ignore"):
        R = (A.astype(np.float64)[:, :, None] * B.astype(np.float64)[None, :, :]).sum(axis=1)
    fa, fb = np.where(np.isfinite(A), np.abs(A), 0.0), np.where(np.isfinite(B), np.abs(B), 0.0)
    S = fa.astype(np.float64) @ fb.astype(np.float64)
    return A, B, R, S


def subnormal_inputs(N=256, seed=5):
    """Uniform values with 4000 entries replaced by subnormal and tiny normal
    magnitudes (down to the smallest subnormal, 1.4e-45), whose bf16 lo parts
    are subnormal too."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (N, N)).astype(np.float32)
    mags = np.array([1e-39, 2.5e-40, 7e-45, 1.4e-45, 1.17549435e-38, 3e-38, 1e-37, 5e-37, 2e-36],
                    np.float32)
    idx = rng.integers(0, N * N, 4000)
    sign = np.sign(rng.uniform(-1, 1, 4000)).astype(np.float32)
    x.reshape(-1)[idx] = mags[rng.integers(0, len(mags), 4000)] * sign
    return x


# --- generators ---------------------------------------------------------------------

def test_fill_pattern_matches_reference_init():
    n = 2 * 512 * 512
    a, b = dev_pattern(512, 2)
    ea, eb = oracle.pattern_init(n)
    assert np.array_equal(a.cpu().numpy(), ea) and np.array_equal(b.cpu().numpy(), eb)
    a2 = dev_empty(1000); b2 = dev_empty(1000)
    hb.fill_pattern(a2, b2, 1000, first=123457)
    ea, eb = oracle.pattern_init(1000, first=123457)
    assert np.array_equal(a2.cpu().numpy(), ea) and np.array_equal(b2.cpu().numpy(), eb)


def test_fill_random_matches_oracle_generators():
    n = 1 << 18
    x = dev_empty(n)
    hb.fill_random(x, n, hb.GEN_UNIFORM, 2024, 77)
    assert np.array_equal(x.cpu().numpy(), oracle.gen("uniform", 2024, n, 77))
    for kind, name in [(hb.GEN_LOGUNIFORM, "loguniform"), (hb.GEN_NORMAL_OUTLIER, "normal_outlier")]:
        hb.fill_random(x, n, kind, 99, 5)
        g = x.cpu().numpy(); o = oracle.gen(name, 99, n, 5)
        # device and host libm may differ by an ulp in double; after rounding
        # to float that leaves at most a 1-ulp difference on rare elements
        ulps = np.abs(g.view(np.int32).astype(np.int64) - o.view(np.int32).astype(np.int64))
        assert ulps.max() <= 1 and (ulps > 0).mean() < 1e-3, (name, ulps.max(), (ulps > 0).mean())


# --- the reference configuration: N_512__NITER_2 ----------------------------------------

def test_sgemm_reference_pattern_N512_NITER2_bit_exact():
    N, NITER = 512, 2
    a, b = dev_pattern(N, NITER)
    c = dev_empty(NITER * N * N)
    hb.sgemm(a, b, c, N, NITER)
    torch.cuda.synchronize()
    got = c.cpu().numpy().reshape(NITER, N, N)
    ha, hbm = oracle.pattern_init(NITER * N * N)
    for it in range(NITER):
        sl = slice(it * N * N, (it + 1) * N * N)
        golden = oracle.host_mm(ha[sl].reshape(N, N), hbm[sl].reshape(N, N))  # main.cpp:54-56
        assert np.array_equal(golden, oracle.pattern_slice(N, it))
        sse, ndiff = oracle.sse(golden, got[it])                             # main.cpp:120-134
        assert sse == 0.0 and ndiff == 0
        assert np.array_equal(got[it], golden)


def test_sgemm_uniform_N512_NITER2_within_fp32_bound():
    N, NITER = 512, 2
    a = dev_uniform(NITER * N * N, 1); b = dev_uniform(NITER * N * N, 2)
    c = dev_empty(NITER * N * N)
    hb.sgemm(a, b, c, N, NITER)
    torch.cuda.synchronize()
    A = a.cpu().numpy().reshape(NITER, N, N); B = b.cpu().numpy().reshape(NITER, N, N)
    C = c.cpu().numpy().reshape(NITER, N, N)
    for it in range(NITER):
        R, S = oracle.gemm_f64(A[it], B[it])
        err = np.abs(C[it].astype(np.float64) - R)
        assert np.all(err <= gamma(N + 2) * S)
        e = oracle.errors(C[it], R, S)
        assert e["e_cw"] <= 1e-5 * np.sqrt(N)
        golden = oracle.host_mm(A[it], B[it])
        sse, _ = oracle.sse(golden, C[it])
        assert sse < 0.01


@pytest.mark.parametrize("N,niter", [(1, 1), (7, 2), (16, 1), (100, 3), (129, 1), (300, 2), (640, 1)])
def test_sgemm_ragged_sizes(N, niter):
    a, b = dev_pattern(N, niter)
    c = dev_empty(niter * N * N)
    hb.sgemm(a, b, c, N, niter)
    torch.cuda.synchronize()
    got = c.cpu().numpy().reshape(niter, N, N)
    for it in range(niter):
        assert np.array_equal(got[it], oracle.pattern_slice(N, it)), (N, it)
    a = dev_uniform(niter * N * N, 3); b = dev_uniform(niter * N * N, 4)
    hb.sgemm(a, b, c, N, niter)
    torch.cuda.synchronize()
    A = a.cpu().numpy().reshape(niter, N, N); B = b.cpu().numpy().reshape(niter, N, N)
    C = c.cpu().numpy().reshape(niter, N, N)
    for it in range(niter):
        R, S = oracle.gemm_f64(A[it], B[it])
        assert np.all(np.abs(C[it] - R) <= gamma(N + 2) * S), (N, it)


# (300, 260, 256) and (2400, 2100, 32 or 320) take the clamped-load tiles
# (ragged M and N, K a multiple of the stage depth, aligned strides): the
# 64 x 64 and the 256 x 256 x 32 configurations respectively.  In the 256 x 256
# configuration, waves whose sub-tile lies outside M x N skip their MFMAs:
# (4160, 1300, 96) has a bottom edge of 64 rows and a right edge of 20
# columns, so its corner tiles run one wave of eight.
@pytest.mark.parametrize("M,N,K,pad", [(200, 136, 72, (5, 3, 1)), (256, 256, 256, (4, 4, 4)),
                                       (64, 512, 1000, (0, 0, 0)), (513, 65, 33, (2, 0, 7)),
                                       (300, 260, 256, (4, 4, 4)), (2400, 2100, 32, (4, 4, 8)),
                                       (2400, 2100, 320, (4, 4, 8)), (4160, 1300, 96, (0, 0, 0))])
def test_sgemm_strided_rectangular(M, N, K, pad):
    batch = 3
    lda, ldb, ldc = K + pad[0], N + pad[1], N + pad[2]
    sA, sB, sC = M * lda + 8, K * ldb + 4, M * ldc + 12
    a = dev_uniform(batch * sA, 5); b = dev_uniform(batch * sB, 6)
    c = torch.full((batch * sC,), -7.0, dtype=torch.float32, device="cuda")
    hb.sgemm_strided(a, b, c, M, N, K, lda, ldb, ldc, sA, sB, sC, batch)
    torch.cuda.synchronize()
    Ah, Bh, Ch = a.cpu().numpy(), b.cpu().numpy(), c.cpu().numpy()
    written = np.zeros(batch * sC, bool)
    for i in range(batch):
        A = np.lib.stride_tricks.as_strided(Ah[i * sA:], (M, K), (lda * 4, 4)).copy()
        B = np.lib.stride_tricks.as_strided(Bh[i * sB:], (K, N), (ldb * 4, 4)).copy()
        C = np.lib.stride_tricks.as_strided(Ch[i * sC:], (M, N), (ldc * 4, 4)).copy()
        R, S = oracle.gemm_f64(A, B)
        assert np.all(np.abs(C - R) <= gamma(K + 2) * S)
        for r in range(M):
            written[i * sC + r * ldc: i * sC + r * ldc + N] = True
    assert np.all(Ch[~written] == -7.0)  # padding and gaps untouched


def test_sgemm_special_values_follow_ieee():
    """NaN and infinities propagate as in IEEE arithmetic: every NaN, +Inf and
    -Inf of the float64 product is in the same place, and the finite elements
    are within the FP32 bound."""
    A, B, R, S = special_inputs()
    M, K = A.shape; N = B.shape[1]
    a = torch.from_numpy(A.ravel()).cuda(); b = torch.from_numpy(B.ravel()).cuda()
    c = dev_empty(M * N)
    hb.sgemm_strided(a, b, c, M, N, K, K, N, N)
    torch.cuda.synchronize()
    C = c.cpu().numpy().reshape(M, N)
    assert np.isnan(R).sum() > 100 and np.isposinf(R).any() and np.isneginf(R).any()
    assert np.array_equal(np.isnan(C), np.isnan(R))
    assert np.array_equal(np.isposinf(C), np.isposinf(R))
    assert np.array_equal(np.isneginf(C), np.isneginf(R))
    fin = np.isfinite(R)
    assert np.all(np.abs(C[fin] - R[fin]) <= gamma(K + 2) * S[fin])


def test_sgemm_keeps_subnormals():
    """C = X . I returns X bit for bit, subnormal entries included: the FP32
    matrix cores neither flush subnormal inputs nor subnormal products."""
    x = subnormal_inputs()
    N = x.shape[0]
    assert ((np.abs(x) < 2.0 ** -126) & (x != 0)).sum() > 1000
    xs = torch.from_numpy(x.ravel()).cuda()
    eye = torch.eye(N, dtype=torch.float32, device="cuda").reshape(-1).contiguous()
    c = dev_empty(N * N)
    hb.sgemm(xs, eye, c, N, 1)
    torch.cuda.synchronize()
    assert np.array_equal(c.cpu().numpy().reshape(N, N), x)


def test_verify_f64_matches_oracle_metrics():
    M, N, K = 256, 192, 320
    a = dev_uniform(M * K, 7); b = dev_uniform(K * N, 8)
    c = dev_empty(M * N)
    hb.sgemm_strided(a, b, c, M, N, K, K, N, N)
    c[5] += 0.125  # a planted error must be seen by both
    st = hb.verify_f64(a, b, c, M, N, K)
    A = a.cpu().numpy().reshape(M, K); B = b.cpu().numpy().reshape(K, N)
    C = c.cpu().numpy().reshape(M, N)
    R, S = oracle.gemm_f64(A, B)
    e = oracle.errors(C, R, S)
    for k in ["e_cw", "e_nw", "e_F", "max_abs_err", "max_abs_ref"]:
        assert st[k] == pytest.approx(e[k], rel=1e-9), k
    assert st["max_abs_err"] >= 0.12 and st["nonfinite"] == 0
    part = hb.verify_f64(a, b, c, M, N, K, row_begin=64, row_end=128)
    e2 = oracle.errors(C[64:128], R[64:128], S[64:128])
    assert part["e_cw"] == pytest.approx(e2["e_cw"], rel=1e-9)


# --- full size: N = 8192, NITER = 10 -------------------------------------------------

def test_sgemm_full_size_pattern_bit_exact_and_uniform_sampled():
    N, NITER = 8192, 10
    a, b = dev_pattern(N, NITER)
    c = dev_empty(NITER * N * N)
    hb.sgemm(a, b, c, N, NITER)
    torch.cuda.synchronize()
    C = c.view(NITER, N, N)
    for it in range(NITER):
        expect = pattern_closed_form_gpu(N, it)
        assert torch.equal(C[it], expect), it
    for (it, y, x) in [(0, 0, 0), (9, 8191, 8191), (4, 1234, 5678)]:
        assert C[it, y, x].item() == oracle.pattern_value(N, it, y, x)
    del a, b
    # uniform inputs: FP64 oracle on a row sample of the last slice
    a = dev_uniform(NITER * N * N, 11); b = dev_uniform(NITER * N * N, 12)
    hb.sgemm(a, b, c, N, NITER)
    torch.cuda.synchronize()
    it, rows = NITER - 1, slice(4096, 4096 + 32)
    A = a.view(NITER, N, N)[it, rows].cpu().numpy()
    B = b.view(NITER, N, N)[it].cpu().numpy()
    R, S = oracle.gemm_f64(A, B)
    Cs = C[it, rows].cpu().numpy()
    assert np.all(np.abs(Cs - R) <= gamma(N + 2) * S)
    st = hb.verify_f64(a.view(NITER, N, N)[it], b.view(NITER, N, N)[it], C[it], N, N, N,
                       row_begin=4096, row_end=4096 + 32)
    assert st["e_cw"] == pytest.approx(oracle.errors(Cs, R, S)["e_cw"], rel=1e-9)
    assert st["e_cw"] <= 1e-5 * np.sqrt(N)
