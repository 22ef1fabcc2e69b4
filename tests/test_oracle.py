"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the restatement (a fused
multiply-add instead of separate rounding, a double accumulator, a wrong
residue, a transposed operand, a wrong rounding mode) fails it.  The reference
holds no golden vectors for apps/sgemm; its only known-answer test is the
integer pattern of apps/sgemm/main.cpp:46-52 with the SSE check of :120-134.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle

U = 2.0 ** -24


def _np_host_mm(A, B):
    """main.cpp:15-26 in numpy float32 scalars: separate rounding of * and +."""
    N = A.shape[0]
    C = np.zeros((N, N), np.float32)
    for y in range(N):
        for x in range(N):
            s = np.float32(0.0)
            for z in range(N):
                s = np.float32(s + np.float32(A[y, z] * B[z, x]))
            C[y, x] = s
    return C


def _round_f32(fr):
    """Correctly rounded (nearest-even) float32 of an exact Fraction."""
    x = np.float32(float(fr))
    best = None
    for c in (np.nextafter(x, np.float32(-np.inf)), x, np.nextafter(x, np.float32(np.inf))):
        d = abs(Fraction(float(c)) - fr)
        if best is None or d < best[0] or (d == best[0] and (int(c.view(np.uint32)) & 1) == 0):
            best = (d, c)
    return np.float32(best[1])


def _exact_fmaf_mm(A, B):
    N = A.shape[0]
    C = np.zeros((N, N), np.float32)
    for y in range(N):
        for x in range(N):
            s = np.float32(0.0)
            for z in range(N):
                s = _round_f32(Fraction(float(A[y, z])) * Fraction(float(B[z, x])) + Fraction(float(s)))
            C[y, x] = s
    return C


# --- generators ----------------------------------------------------------------------

def test_splitmix64_published_vector():
    # SplitMix64 seeded with 1234567: first five outputs published with the
    # algorithm's reference implementation (Vigna, splitmix64.c test vector).
    expect = [6457827717110365317, 3203168211198807973, 9817491932198370423,
              4593380528125082431, 16408922859458223821]
    assert [oracle.rng64(1234567, i) for i in range(5)] == expect


def _np_rng64(seed, idx):
    g = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * g
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def test_uniform_generator_formula():
    n, seed, first = 4096, 42, 1000
    x = oracle.gen("uniform", seed, n, first)
    r = _np_rng64(seed, np.arange(first, first + n))
    u = (r >> np.uint64(40)).astype(np.float64) * 2.0 ** -24
    assert np.array_equal(x, (2 * u - 1).astype(np.float32))
    assert x.min() >= -1.0 and x.max() < 1.0
    # slices are independent of how the stream is cut
    assert np.array_equal(oracle.gen("uniform", seed, 100, first + 7), x[7:107])


def test_loguniform_generator_distribution():
    x = oracle.gen("loguniform", 7, 200000)
    lg = np.log10(np.abs(x.astype(np.float64)))
    assert lg.min() >= -3.0 - 1e-6 and lg.max() <= 3.0 + 1e-6
    # uniform in log10 |x|: deciles near -3 + 0.6 k
    q = np.quantile(lg, np.linspace(0.1, 0.9, 9))
    assert np.allclose(q, -3 + 6 * np.linspace(0.1, 0.9, 9), atol=0.03)
    assert abs((x > 0).mean() - 0.5) < 0.01


def test_normal_outlier_generator_distribution():
    n, seed = 300000, 11
    x = oracle.gen("normal_outlier", seed, n).astype(np.float64)
    r3 = _np_rng64(seed, 3 * np.arange(n) + 2)
    mask = (r3 >> np.uint64(11)).astype(np.float64) * 2.0 ** -53 < 0.01
    assert abs(mask.mean() - 0.01) < 0.002
    z = np.where(mask, x / 1000.0, x)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
    assert abs(np.mean(np.abs(z) < 1.0) - 0.6827) < 0.005
    assert abs(np.mean(np.abs(z) < 2.0) - 0.9545) < 0.003


# --- apps/sgemm/main.cpp:15-26 (host_mm) ----------------------------------------

def test_host_mm_matches_float32_emulation_bit_exact():
    rng = np.random.default_rng(3)
    A = rng.uniform(-1, 1, (7, 7)).astype(np.float32)
    B = rng.uniform(-1, 1, (7, 7)).astype(np.float32)
    assert np.array_equal(oracle.host_mm(A, B), _np_host_mm(A, B))


def test_host_mm_rounds_products_separately():
    # With separate rounding host_mm differs from the fmaf chain on generic
    # inputs; a restatement that fused them (or accumulated in double) would not.
    rng = np.random.default_rng(5)
    A = rng.uniform(-1, 1, (32, 32)).astype(np.float32)
    B = rng.uniform(-1, 1, (32, 32)).astype(np.float32)
    h = oracle.host_mm(A, B)
    f = oracle.fmaf_chain_mm(A, B)
    d = A.astype(np.float64) @ B.astype(np.float64)
    assert not np.array_equal(h, f)
    assert not np.array_equal(h, d.astype(np.float32))


def test_fmaf_chain_matches_exact_fma_emulation():
    rng = np.random.default_rng(9)
    A = rng.uniform(-1, 1, (5, 5)).astype(np.float32)
    B = rng.uniform(-1, 1, (5, 5)).astype(np.float32)
    assert np.array_equal(oracle.fmaf_chain_mm(A, B), _exact_fmaf_mm(A, B))


def test_host_mm_rows_and_omp_are_bit_identical():
    rng = np.random.default_rng(1)
    A = rng.uniform(-1, 1, (64, 64)).astype(np.float32)
    B = rng.uniform(-1, 1, (64, 64)).astype(np.float32)
    full = oracle.host_mm(A, B)
    part = oracle.host_mm_rows(A, B, 10, 30)
    assert np.array_equal(part[10:30], full[10:30]) and not part[:10].any()
    omp = oracle.host_mm_rows(A, B, 0, 64, threads=4)
    assert np.array_equal(omp, full)


def test_host_mm_within_gamma_bound_of_fp64():
    rng = np.random.default_rng(2)
    N = 96
    A = rng.uniform(-1, 1, (N, N)).astype(np.float32)
    B = rng.uniform(-1, 1, (N, N)).astype(np.float32)
    C = oracle.host_mm(A, B)
    R, S = oracle.gemm_f64(A, B)
    gamma = N * U / (1 - N * U)
    assert np.all(np.abs(C - R) <= gamma * S)


# --- apps/sgemm/main.cpp:46-52 pattern and its closed form ----------------------

def _brute_pattern(N, it):
    o = it * N * N
    y = np.arange(N)[:, None]; z = np.arange(N)[None, :]
    a = (o + N * y + z) % 7                  # mat1[o + N*y + z]
    b = (o + N * z.T + np.arange(N)[None, :]) % 3   # mat2[o + N*z + x]
    return a.astype(np.int64) @ b.astype(np.int64)


@pytest.mark.parametrize("N,it", [(16, 0), (16, 1), (21, 3), (22, 0), (37, 2), (64, 1)])
def test_pattern_closed_form_matches_brute_force(N, it):
    C = oracle.pattern_slice(N, it).astype(np.int64)
    assert np.array_equal(C, _brute_pattern(N, it))


def test_pattern_init_matches_reference_loop():
    a, b = oracle.pattern_init(1000, first=12345)
    i = np.arange(12345, 13345)
    assert np.array_equal(a, (i % 7).astype(np.float32))
    assert np.array_equal(b, (i % 3).astype(np.float32))


def test_host_mm_on_pattern_is_exact_at_reference_size():
    # N_512__NITER_2 (apps/sgemm/tests.mk:1): every partial sum is an integer
    # < 12 N < 2^24, so host_mm equals the closed form bit for bit.
    N, NITER = 512, 2
    a, b = oracle.pattern_init(NITER * N * N)
    for it in range(NITER):
        sl = slice(it * N * N, (it + 1) * N * N)
        C = oracle.host_mm(a[sl].reshape(N, N), b[sl].reshape(N, N))
        assert np.array_equal(C, oracle.pattern_slice(N, it))


def test_pattern_value_at_large_n_matches_int_sum():
    N = 8192
    for (it, y, x) in [(0, 0, 0), (9, 8191, 8191), (3, 100, 4000)]:
        o = it * N * N
        z = np.arange(N, dtype=np.int64)
        ref = int((((o + N * y + z) % 7) * ((o + N * z + x) % 3)).sum())
        assert oracle.pattern_value(N, it, y, x) == ref
        assert ref < 2 ** 24


# --- apps/sgemm/main.cpp:120-134 (SSE check) -------------------------------------

def test_sse_check_accepts_equal_and_catches_planted_error():
    e = np.arange(1000, dtype=np.float32)
    s, nd = oracle.sse(e, e.copy())
    assert s == 0.0 and nd == 0
    bad = e.copy(); bad[17] += 0.25
    s, nd = oracle.sse(e, bad)
    assert nd == 1 and s == np.float32(0.0625)
    bad[500] += 1.0
    s, nd = oracle.sse(e, bad)
    assert nd == 2 and s >= 0.01  # the reference FAIL threshold (main.cpp:131)


# --- FP64 oracle and error metrics ---------------------------------------------------

def test_gemm_f64_matches_exact_integer_and_numpy():
    rng = np.random.default_rng(4)
    A = rng.integers(-50, 50, (33, 45)).astype(np.float32)
    B = rng.integers(-50, 50, (45, 29)).astype(np.float32)
    R, S = oracle.gemm_f64(A, B)
    assert np.array_equal(R, A.astype(np.int64) @ B.astype(np.int64))
    assert np.array_equal(S, np.abs(A).astype(np.int64) @ np.abs(B).astype(np.int64))
    Af = rng.uniform(-1, 1, (40, 50)).astype(np.float32)
    Bf = rng.uniform(-1, 1, (50, 30)).astype(np.float32)
    R, _ = oracle.gemm_f64(Af, Bf)
    assert np.allclose(R, Af.astype(np.float64) @ Bf.astype(np.float64), rtol=1e-13, atol=1e-13)


def test_error_metrics_hand_computed():
    R = np.array([[1.0, -2.0], [4.0, 0.0]])
    S = np.array([[2.0, 2.0], [8.0, 1.0]])
    C = np.array([[1.5, -2.0], [4.0, 0.25]], np.float32)
    e = oracle.errors(C, R, S)
    assert e["e_cw"] == pytest.approx(0.25)          # max(0.5/2, 0.25/1)
    assert e["max_abs_err"] == pytest.approx(0.5)
    assert e["e_nw"] == pytest.approx(0.5 / 4.0)
    assert e["e_F"] == pytest.approx(np.sqrt((0.25 + 0.0625) / 21.0))
    C2 = C.copy(); C2[0, 0] = np.nan
    assert oracle.errors(C2, R, S)["nonfinite"] == 1


# --- BF16 split --------------------------------------------------------------------

def test_bf16_rne_matches_torch():
    torch = pytest.importorskip("torch")
    x = np.concatenate([oracle.gen("loguniform", 3, 20000), oracle.gen("uniform", 4, 20000),
                        np.array([0.0, -0.0, 1.0, 3.0e38, 1e-40], np.float32)])
    hi, _ = oracle.split_bf16(x)
    t = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(hi, t)


def test_bf16_split_reconstruction_bound():
    x = oracle.gen("loguniform", 8, 50000)
    hi, lo = oracle.split_bf16(x)
    h = oracle.bf16_bits_to_f32(hi).astype(np.float64)
    l = oracle.bf16_bits_to_f32(lo).astype(np.float64)
    xd = x.astype(np.float64)
    # bf16 keeps 8 significant bits: unit roundoff 2^-8 for hi, 2^-16 for hi + lo
    assert np.all(np.abs(xd - h) <= 2.0 ** -8 * np.abs(xd))
    assert np.all(np.abs(xd - h - l) <= 2.0 ** -16 * np.abs(xd))
    assert np.mean(np.abs(xd - h) > 2.0 ** -10 * np.abs(xd)) > 0.2  # bound is not vacuous
    # lo is exactly zero for values already representable in bf16 (0..6 pattern)
    hi2, lo2 = oracle.split_bf16(np.arange(7, dtype=np.float32))
    assert not lo2.any() and np.array_equal(oracle.bf16_bits_to_f32(hi2), np.arange(7, dtype=np.float32))


# --- FP8 e4m3 -------------------------------------------------------------------------

def test_fp8_codes_roundtrip():
    for b in range(256):
        if b & 0x7F == 0x7F:
            continue  # NaN encodings of e4m3fn
        v = oracle.fp8_decode(b)
        assert oracle.fp8_encode(v) == b or (v == 0.0 and b in (0x00, 0x80))
    assert oracle.fp8_decode(0x7E) == 448.0 and oracle.fp8_decode(0x01) == 2.0 ** -9


def test_fp8_round_matches_torch():
    torch = pytest.importorskip("torch")
    x = np.concatenate([oracle.gen("uniform", 5, 20000) * 448.0,
                        oracle.gen("loguniform", 6, 20000) * 0.4]).astype(np.float32)
    x = x[np.abs(x) <= 448.0]
    mine = np.array([oracle.fp8_round(v) for v in x], np.float32)
    t = torch.from_numpy(x).to(torch.float8_e4m3fn).to(torch.float32).numpy()
    assert np.array_equal(mine, t)


def test_fp8_tile_exponent_bounds():
    for amax in [1e-3, 0.5, 1.0, 3.0, 224.0, 224.1, 448.0, 449.0, 1000.0, 3.7e3]:
        e = oracle.fp8_tile_exponent(amax)
        r = amax / 2.0 ** e
        assert 224.0 < r <= 448.0, (amax, e, r)
    assert oracle.fp8_tile_exponent(0.0) == 0


def test_fp8_quantize_pattern_is_exact_and_residual_bound():
    # integer pattern values 0..6: power-of-two scales keep them exact, residual 0
    a, _ = oracle.pattern_init(256 * 256)
    X = a.reshape(256, 256)
    Xq, Xr, e = oracle.fp8_quantize(X)
    assert np.array_equal(Xq, X) and not Xr.any() and e.shape == (2, 2)
    X = oracle.gen("normal_outlier", 2, 256 * 384).reshape(256, 384)
    Xq, Xr, e = oracle.fp8_quantize(X)
    Xd = X.astype(np.float64)
    # residual is bf16 of (X - Xq): X - Xq - Xr within 2^-8 of |X - Xq|
    d = Xd - Xq.astype(np.float64)
    assert np.all(np.abs(d - Xr) <= 2.0 ** -8 * np.abs(d))
    for bi in range(2):
        for bj in range(3):
            t = X[bi * 128:(bi + 1) * 128, bj * 128:(bj + 1) * 128]
            assert 224.0 < np.abs(t).max() / 2.0 ** e[bi, bj] <= 448.0


def test_fp8_tile_scale_ignores_non_finite_entries():
    """A NaN or an infinity in a 128 x 128 tile leaves its exponent at what
    the tile's finite entries give, and every finite entry's q and r at what
    they are with the non-finite entries replaced by zero."""
    rng = np.random.default_rng(8)
    X = rng.normal(0, 3, (256, 256)).astype(np.float32)
    Y = X.copy()
    Y[5, 30] = np.inf; Y[140, 200] = -np.inf; Y[70, 150] = np.nan
    Z = Y.copy()
    Z[~np.isfinite(Z)] = 0.0
    qy, ry, ey = oracle.fp8_quantize(Y)
    qz, rz, ez = oracle.fp8_quantize(Z)
    assert np.array_equal(ey, ez)
    fin = np.isfinite(Y)
    assert np.array_equal(qy[fin], qz[fin]) and np.array_equal(ry[fin], rz[fin])
    # an infinity alone in its tile would otherwise push the scale to 2^120
    assert ez.max() < 10
