"""GPU parity of the BF16x3 split-precision SGEMM, through the C-ABI.

No reference function exists for this kernel; it is pinned by
 * bit-exactness on the reference integer pattern (apps/sgemm/main.cpp:46-52):
   values 0..6 are exact in bf16, so lo = 0 and every product is exact;
 * a rigorous per-element bound against the FP64 oracle: the split drops at
   most 3.1 * 2^-16 |a||b| per product (lo.lo and the rounding of lo), and the
   3K products are summed in FP32: |C - R| <= (3.1 * 2^-16 + gamma_{3K+2}) S;
 * the FP32-level accuracy claim at N = 8192 on log-uniform inputs: max
   componentwise error no worse than 2x the plain FP32 kernel's on the same
   inputs.
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


def dev_fill(n, kind, seed):
    x = dev_empty(n)
    hb.fill_random(x, n, kind, seed)
    return x


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


def test_bf16x3_split_keeps_subnormals():
    """With B = I, C is the oracle's hi + lo of A bit for bit, also where A or
    its lo part is subnormal: neither the hardware bf16 conversion nor the
    BF16 matrix cores flush them."""
    x = subnormal_inputs()
    N = x.shape[0]
    xs = torch.from_numpy(x.ravel()).cuda()
    eye = torch.eye(N, dtype=torch.float32, device="cuda").reshape(-1).contiguous()
    c = torch.empty(N * N, dtype=torch.float32, device="cuda")
    hb.sgemm_bf16x3(xs, eye, c, N, 1)
    torch.cuda.synchronize()
    hi, lo = oracle.split_bf16(x)
    lo_f = oracle.bf16_bits_to_f32(lo)
    assert ((np.abs(lo_f) < 2.0 ** -126) & (lo_f != 0)).sum() > 100
    want = oracle.bf16_bits_to_f32(hi) + lo_f
    assert np.array_equal(c.cpu().numpy().reshape(N, N), want)


def test_bf16x3_bf16_overflow_threshold():
    """Finite inputs below the bf16 overflow threshold (2^128 - 2^119, about
    3.3962e38) come through finite; at or above it, bf16 rounding of x is an
    infinity and the result is NaN (one value per row, B = I)."""
    vals = np.array([3.38e38, 3.3895e38, 3.395e38, 3.3962e38, 3.4e38, np.finfo(np.float32).max,
                     -3.4e38, 1.0], np.float32)
    N = 64
    x = np.zeros((N, N), np.float32)
    x[np.arange(len(vals)), 0] = vals
    eye = torch.eye(N, dtype=torch.float32, device="cuda").reshape(-1).contiguous()
    c = torch.empty(N * N, dtype=torch.float32, device="cuda")
    hb.sgemm_bf16x3(torch.from_numpy(x.ravel()).cuda(), eye, c, N, 1)
    torch.cuda.synchronize()
    got = c.cpu().numpy().reshape(N, N)[np.arange(len(vals)), 0]
    below = np.abs(vals) < 2.0 ** 128 - 2.0 ** 119
    assert np.all(np.isfinite(got[below]))
    assert np.all(np.abs(got[below] - vals[below]) <= 2.0 ** -15 * np.abs(vals[below]))
    assert np.all(np.isnan(got[~below]))


def test_bf16x3_special_values_stay_non_finite():
    """A non-finite input makes its results non-finite, as in IEEE arithmetic,
    and every NaN of the float64 product is NaN here.  An infinity may come
    out as NaN: its lo part is not finite, and Inf x 0 cross terms are NaN.
    Finite elements are within the split bound."""
    A, B, R, S = special_inputs()
    M, K = A.shape; N = B.shape[1]
    a = torch.from_numpy(A.ravel()).cuda(); b = torch.from_numpy(B.ravel()).cuda()
    c = torch.empty(M * N, dtype=torch.float32, device="cuda")
    hb.gemm_bf16x3(a, b, c, M, N, K)
    torch.cuda.synchronize()
    C = c.cpu().numpy().reshape(M, N)
    assert np.array_equal(np.isfinite(C), np.isfinite(R))
    assert np.all(np.isnan(C[np.isnan(R)]))
    fin = np.isfinite(R)
    assert np.all(np.abs(C[fin] - R[fin]) <= (3.1 * 2.0 ** -16 + gamma(3 * K + 2)) * S[fin])


def test_bf16x3_reference_pattern_N512_NITER2_bit_exact():
    N, NITER = 512, 2
    a = dev_empty(NITER * N * N); b = dev_empty(NITER * N * N); c = dev_empty(NITER * N * N)
    hb.fill_pattern(a, b, NITER * N * N)
    hb.sgemm_bf16x3(a, b, c, N, NITER)
    torch.cuda.synchronize()
    got = c.cpu().numpy().reshape(NITER, N, N)
    for it in range(NITER):
        sse, nd = oracle.sse(oracle.pattern_slice(N, it), got[it])
        assert sse == 0.0 and nd == 0


@pytest.mark.parametrize("N,niter,kind", [(512, 2, hb.GEN_UNIFORM), (256, 1, hb.GEN_LOGUNIFORM),
                                          (1, 1, hb.GEN_UNIFORM), (77, 2, hb.GEN_UNIFORM),
                                          (300, 1, hb.GEN_LOGUNIFORM), (520, 1, hb.GEN_UNIFORM)])
def test_bf16x3_within_split_bound(N, niter, kind):
    a = dev_fill(niter * N * N, kind, 21); b = dev_fill(niter * N * N, kind, 22)
    c = dev_empty(niter * N * N)
    hb.sgemm_bf16x3(a, b, c, N, niter)
    torch.cuda.synchronize()
    A = a.cpu().numpy().reshape(niter, N, N); B = b.cpu().numpy().reshape(niter, N, N)
    C = c.cpu().numpy().reshape(niter, N, N)
    for it in range(niter):
        R, S = oracle.gemm_f64(A[it], B[it])
        bound = (3.1 * 2.0 ** -16 + gamma(3 * N + 2)) * S
        assert np.all(np.abs(C[it] - R) <= bound), (N, it)
        e = oracle.errors(C[it], R, S)
        assert e["e_cw"] <= 1e-5 * np.sqrt(N)


@pytest.mark.parametrize("N", [300, 512, 1536])
def test_bf16x3_identity_reproduces_oracle_split(N):
    """With B = I (and with A = I) each output is lo + 0 + hi of the other
    operand's element, and hi + lo is exact in FP32, so C equals the oracle's
    split summed in float32 bit for bit: this pins the in-register split of
    the A path and of the B path (N = 300: ragged 64 x 64 tiles; 512: interior
    64 x 64 tiles; 1536: 128 x 128 tiles, the grid being under 256 tiles of
    256 x 256)."""
    x = dev_fill(N * N, hb.GEN_LOGUNIFORM, 41)
    eye = torch.eye(N, dtype=torch.float32, device="cuda").reshape(-1).contiguous()
    c1 = dev_empty(N * N); c2 = dev_empty(N * N)
    hb.sgemm_bf16x3(x, eye, c1, N, 1)
    hb.sgemm_bf16x3(eye, x, c2, N, 1)
    torch.cuda.synchronize()
    xh = x.cpu().numpy().reshape(N, N)
    hi, lo = oracle.split_bf16(xh)
    want = oracle.bf16_bits_to_f32(hi) + oracle.bf16_bits_to_f32(lo)
    assert want.dtype == np.float32 and (want != xh).any()  # the split drops bits
    assert np.array_equal(c1.cpu().numpy().reshape(N, N), want)
    assert np.array_equal(c2.cpu().numpy().reshape(N, N), want)


@pytest.mark.parametrize("M,N,K,lda,ldb,ldc", [(300, 260, 256, 256, 260, 264),
                                                (1000, 772, 64, 68, 776, 772),
                                                (257, 513, 96, 97, 513, 515)])
def test_gemm_bf16x3_rectangular_within_split_bound(M, N, K, lda, ldb, ldc):
    """Rectangular entry with leading dimensions: ragged M and N with K a
    multiple of 32 and A 16-byte loadable take the clamped-load tiles (the
    first two shapes), an odd lda the predicated ones; elements outside M x N
    stay untouched."""
    a = dev_fill(M * lda, hb.GEN_LOGUNIFORM, 51); b = dev_fill(K * ldb, hb.GEN_UNIFORM, 52)
    c = torch.full((M * ldc,), -3.0, dtype=torch.float32, device="cuda")
    hb.gemm_bf16x3(a, b, c, M, N, K, lda=lda, ldb=ldb, ldc=ldc)
    torch.cuda.synchronize()
    A = a.cpu().numpy().reshape(M, lda)[:, :K]
    B = b.cpu().numpy().reshape(K, ldb)[:, :N]
    Cf = c.cpu().numpy().reshape(M, ldc)
    R, S = oracle.gemm_f64(np.ascontiguousarray(A), np.ascontiguousarray(B))
    bound = (3.1 * 2.0 ** -16 + gamma(3 * K + 2)) * S
    assert np.all(np.abs(Cf[:, :N] - R) <= bound)
    assert np.all(Cf[:, N:] == -3.0)


def test_bf16x3_pattern_ragged_sizes():
    for N in (3, 100, 257):
        a = dev_empty(N * N); b = dev_empty(N * N); c = dev_empty(N * N)
        hb.fill_pattern(a, b, N * N)
        hb.sgemm_bf16x3(a, b, c, N, 1)
        torch.cuda.synchronize()
        assert np.array_equal(c.cpu().numpy().reshape(N, N), oracle.pattern_slice(N, 0)), N


def test_bf16x3_deterministic_across_calls():
    for N in (512, 128):
        a = dev_fill(N * N, hb.GEN_UNIFORM, 1); b = dev_fill(N * N, hb.GEN_UNIFORM, 2)
        c1 = dev_empty(N * N); c2 = dev_empty(N * N)
        hb.sgemm_bf16x3(a, b, c1, N, 1)
        hb.sgemm_bf16x3(a, b, c2, N, 1)
        torch.cuda.synchronize()
        assert torch.equal(c1, c2)


def test_bf16x3_fp32_level_error_at_N8192_loguniform():
    """BF16x3 max componentwise error <= 2x the FP32 kernel's, same inputs."""
    N = 8192
    a = dev_fill(N * N, hb.GEN_LOGUNIFORM, 31); b = dev_fill(N * N, hb.GEN_LOGUNIFORM, 32)
    c32 = dev_empty(N * N); c3 = dev_empty(N * N)
    hb.sgemm(a, b, c32, N, 1)
    hb.sgemm_bf16x3(a, b, c3, N, 1)
    torch.cuda.synchronize()
    e32 = hb.verify_f64(a, b, c32, N, N, N)
    e3 = hb.verify_f64(a, b, c3, N, N, N)
    print(f"\nN=8192 log-uniform: e_cw fp32={e32['e_cw']:.3e} bf16x3={e3['e_cw']:.3e} "
          f"e_F fp32={e32['e_F']:.3e} bf16x3={e3['e_F']:.3e}", flush=True)
    assert e32["nonfinite"] == 0 and e3["nonfinite"] == 0
    assert e3["e_cw"] <= 2.0 * e32["e_cw"]


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


def test_bf16x3_full_size_launch_pattern_and_loguniform():
    """The launch the bench times: N = 8192, NITER = 10 slices in one call.
    Reference pattern: every slice bit-exact against the closed form (values
    0..6 and 0..2 are exact in bf16, so lo = 0 and every product is exact).
    Log-uniform inputs (|x| in [1e-3, 1e3], random sign): the last slice of
    the same 10-slice launch within the split bound of the FP64 oracle on a
    row band, and its e_cw over the whole slice no worse than 2x the FP32
    kernel's on the same batch."""
    N, NITER = 8192, 10
    n = NITER * N * N
    a = dev_empty(n); b = dev_empty(n); c = dev_empty(n)
    hb.fill_pattern(a, b, n)
    hb.sgemm_bf16x3(a, b, c, N, NITER)
    torch.cuda.synchronize()
    C = c.view(NITER, N, N)
    for it in range(NITER):
        assert torch.equal(C[it], pattern_closed_form_gpu(N, it)), it
    for (it, y, x) in [(0, 0, 0), (9, 8191, 8191), (6, 4321, 777)]:
        assert C[it, y, x].item() == oracle.pattern_value(N, it, y, x)

    hb.fill_random(a, n, hb.GEN_LOGUNIFORM, 51)
    hb.fill_random(b, n, hb.GEN_LOGUNIFORM, 52)
    hb.sgemm_bf16x3(a, b, c, N, NITER)
    torch.cuda.synchronize()
    it, r0, r1 = NITER - 1, 6000, 6032
    Av, Bv = a.view(NITER, N, N)[it], b.view(NITER, N, N)[it]
    A = Av[r0:r1].cpu().numpy()
    B = Bv.cpu().numpy()
    R, S = oracle.gemm_f64(A, B)
    Cs = C[it, r0:r1].cpu().numpy()
    assert np.all(np.isfinite(Cs))
    assert np.all(np.abs(Cs - R) <= (3.1 * 2.0 ** -16 + gamma(3 * N + 2)) * S)
    band = hb.verify_f64(Av, Bv, C[it], N, N, N, row_begin=r0, row_end=r1)
    assert band["e_cw"] == pytest.approx(oracle.errors(Cs, R, S)["e_cw"], rel=1e-9)
    e3 = hb.verify_f64(Av, Bv, C[it], N, N, N)  # the whole last slice
    # the FP32 kernel on the same 10-slice batch, same last slice
    hb.sgemm(a, b, c, N, NITER)
    torch.cuda.synchronize()
    e32 = hb.verify_f64(Av, Bv, C[it], N, N, N)
    print(f"\nN=8192 NITER=10 last slice: e_cw bf16x3={e3['e_cw']:.3e} fp32={e32['e_cw']:.3e} "
          f"e_nw {e3['e_nw']:.3e}/{e32['e_nw']:.3e} e_F {e3['e_F']:.3e}/{e32['e_F']:.3e}",
          flush=True)
    assert e3["nonfinite"] == 0 and e32["nonfinite"] == 0
    assert e3["e_cw"] <= 2.0 * e32["e_cw"]
