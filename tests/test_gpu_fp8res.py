"""GPU parity of the FP8-e4m3 + BF16-residual GEMM, through the C-ABI.

No reference function exists for this kernel.  It is pinned by
 * the quantiser: the device's e4m3 bytes, q|r bf16 planes and tile exponents
   equal the oracle's (oracle/orc_fp8_quantize, itself pinned against torch's
   float8_e4m3fn and bfloat16 conversions) bit for bit;
 * the product: C within the FP32 accumulation bound of the exact value of
   the algorithm, Aq.Bq + Aq.Br + Ar.Bq, computed in FP64 from the oracle's
   quantised operands: |C - I| <= gamma_{3K+2} (|Aq||Bq| + |Aq||Br| + |Ar||Bq|)
   plus, for the FP8 matrix-core term, 32 * 2^-13 * max|aq bq| per 32-element k
   block.  The FP8 MFMA sums each 32-element block with about 14 significant
   bits relative to the block's largest product, truncating (measured on
   gfx950: 448 + 2^-6 in one block returns 448, in two blocks 448.015625);
   inputs whose magnitudes stay within one binade of each other meet the plain
   FP32 bound, which test_fp8res_without_subnormals_matches_algorithm checks;
 * bit-exactness on the reference's integer pattern (power-of-two tile scales
   keep 0..6 and 0..2 exact in e4m3, so the residual is zero);
 * the error against the true product: |C - AB| <= 2^-8 |Aq||dB| +
   2^-8 |dA||Bq| + |dA||dB| + (the bound above), dX = X - Xq.
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


def rup(x, m):
    return (x + m - 1) // m * m


def planes(ws, M, N, K):
    """Splits the documented workspace layout into its planes (numpy)."""
    Mp, Np, Kp = rup(M, 256), rup(N, 256), rup(K, 128)
    w = ws.cpu().numpy()
    o = 0

    def take(nbytes):
        nonlocal o
        v = w[o:o + nbytes]
        o = rup(o + nbytes, 256)
        return v
    A8 = take(Mp * Kp).reshape(Mp, Kp)
    Ar = take(Mp * Kp * 2).view(np.uint16).reshape(Mp, Kp)
    eA = take((Mp // 128) * (Kp // 128) * 4).view(np.int32).reshape(Mp // 128, Kp // 128)
    B8 = take(Np * Kp).reshape(Np, Kp)
    Br = take(Np * Kp * 2).view(np.uint16).reshape(Np, Kp)
    eB = take((Np // 128) * (Kp // 128) * 4).view(np.int32).reshape(Np // 128, Kp // 128)
    return A8, Ar, eA, B8, Br, eB


def decode_plane(X8, e, rows, cols):
    """e4m3 bytes times their tile's 2^e, as float32 (exact)."""
    lut = np.array([oracle.fp8_decode(b) if (b & 0x7F) != 0x7F else np.nan for b in range(256)])
    scale = np.repeat(np.repeat(2.0 ** e.astype(np.float64), 128, 0), 128, 1)
    return (lut[X8[:rows, :cols]] * scale[:rows, :cols]).astype(np.float32)


def fp8_block_term(Aq, Bq):
    """sum over 32-element k blocks of 32 * 2^-13 * max_k |Aq[i,k] Bq[k,j]|."""
    M, K = Aq.shape
    out = np.zeros((M, Bq.shape[1]))
    for k0 in range(0, K, 32):
        a = np.abs(Aq[:, k0:k0 + 32])[:, :, None]
        b = np.abs(Bq[k0:k0 + 32, :])[None, :, :]
        out += (a * b).max(axis=1)
    return 32 * 2.0 ** -13 * out


def run(M, N, K, kind="normal_outlier", seed=3):
    a = torch.empty(M * K, dtype=torch.float32, device="cuda")
    b = torch.empty(K * N, dtype=torch.float32, device="cuda")
    if kind == "pattern":
        n = max(M * K, K * N)
        ta = torch.empty(n, device="cuda"); tb = torch.empty(n, device="cuda")
        hb.fill_pattern(ta, tb, n)
        a.copy_(ta[:M * K]); b.copy_(tb[:K * N])
    else:
        k = {"normal_outlier": hb.GEN_NORMAL_OUTLIER, "uniform": hb.GEN_UNIFORM}[kind]
        hb.fill_random(a, M * K, k, seed)
        hb.fill_random(b, K * N, k, seed + 1)
    c = torch.empty(M * N, dtype=torch.float32, device="cuda")
    ws = hb.gemm_fp8res(a, b, c, M, N, K)
    torch.cuda.synchronize()
    return (a.cpu().numpy().reshape(M, K), b.cpu().numpy().reshape(K, N),
            c.cpu().numpy().reshape(M, N), ws)


def test_fp8res_pattern_bit_exact():
    M, N, K = 512, 256, 384
    A, B, C, _ = run(M, N, K, "pattern")
    expect = A.astype(np.int64) @ B.astype(np.int64)
    assert np.array_equal(C, expect.astype(np.float32))


@pytest.mark.parametrize("M,N,K", [(256, 128, 256), (300, 200, 100), (512, 384, 640)])
def test_fp8res_quantiser_matches_oracle(M, N, K):
    A, B, C, ws = run(M, N, K)
    A8, Ar_, eA, B8, Br_, eB = planes(ws, M, N, K)
    Aq, Ar, ea = oracle.fp8_quantize(A)
    Bq, Br, eb = oracle.fp8_quantize(np.ascontiguousarray(B.T))
    assert np.array_equal(eA[:ea.shape[0], :ea.shape[1]], ea)
    assert np.array_equal(eB[:eb.shape[0], :eb.shape[1]], eb)
    # residual planes (bf16 bits) and decoded e4m3 planes against the oracle
    assert np.array_equal(oracle.bf16_bits_to_f32(Ar_)[:M, :K], Ar)
    assert np.array_equal(oracle.bf16_bits_to_f32(Br_)[:N, :K], Br)
    assert np.array_equal(decode_plane(A8, eA, M, K), Aq)
    assert np.array_equal(decode_plane(B8, eB, N, K), Bq)
    # padding is zero (the wide kernel reads B's rows up to round_up(N, 256))
    assert not A8[M:].any() and not A8[:, K:].any()
    assert not B8[N:].any() and not B8[:, K:].any() and not Br_[N:].any()


# Widths whose round_up(N, 128) is a multiple of 256 run the 256 x 256
# kernel; at these sizes (few tiles) the others run the 256 x 128 one:
# (256, 512, 384) and (384, 768, 512) take the wide kernel, (300, 200, 100)
# and (300, 700, 200) its predicated-store form; (256, 512, 100) and
# (512, 256, 128) give it the shortest pipeline (Kp = 128: two 64-deep
# stages); (256, 128, 256) and (512, 384, 1024) take the 256 x 128 kernel.
@pytest.mark.parametrize("M,N,K", [(256, 128, 256), (300, 200, 100), (512, 384, 1024),
                                   (256, 512, 384), (384, 768, 512), (300, 700, 200),
                                   (256, 512, 100), (512, 256, 128)])
def test_fp8res_matches_algorithm_within_fp32_bound(M, N, K):
    A, B, C, _ = run(M, N, K, seed=M)
    Aq, Ar, _ = oracle.fp8_quantize(A)
    BqT, BrT, _ = oracle.fp8_quantize(np.ascontiguousarray(B.T))
    Bq, Br = BqT.T.astype(np.float64), BrT.T.astype(np.float64)
    Aq, Ar = Aq.astype(np.float64), Ar.astype(np.float64)
    ideal = Aq @ Bq + Aq @ Br + Ar @ Bq
    absum = np.abs(Aq) @ np.abs(Bq) + np.abs(Aq) @ np.abs(Br) + np.abs(Ar) @ np.abs(Bq)
    hw = fp8_block_term(Aq, Bq)
    err = np.abs(C - ideal)
    assert np.all(err <= gamma(3 * K + 2) * absum + hw)
    print(f"\n{M}x{N}x{K}: max |C - I| / (gamma absum + hw) = "
          f"{(err / (gamma(3 * K + 2) * absum + hw)).max():.3f}", flush=True)
    dA = A.astype(np.float64) - Aq
    dB = B.astype(np.float64) - Bq
    R = A.astype(np.float64) @ B.astype(np.float64)
    algo = (2.0 ** -8 * (np.abs(Aq) @ np.abs(dB)) + 2.0 ** -8 * (np.abs(dA) @ np.abs(Bq))
            + np.abs(dA) @ np.abs(dB))
    assert np.all(np.abs(C - R) <= algo + gamma(3 * K + 2) * absum + hw)


@pytest.mark.parametrize("M,K", [(512, 640), (300, 200)])
def test_fp8res_wide_tile_equals_narrow_tile(M, K):
    """N = 512 runs the 256 x 256 kernel and N = 384 the 256 x 128 one.  B's
    first 384 columns are whole 128-column scaling tiles, so both quantise
    them alike, and each output element sees the same MFMAs in the same order:
    the shared columns of C agree bit for bit."""
    a = torch.empty(M * K, dtype=torch.float32, device="cuda")
    b = torch.empty(K * 512, dtype=torch.float32, device="cuda")
    hb.fill_random(a, M * K, hb.GEN_NORMAL_OUTLIER, 11)
    hb.fill_random(b, K * 512, hb.GEN_NORMAL_OUTLIER, 12)
    b384 = b.view(K, 512)[:, :384].contiguous().view(-1)
    c512 = torch.empty(M * 512, dtype=torch.float32, device="cuda")
    c384 = torch.empty(M * 384, dtype=torch.float32, device="cuda")
    hb.gemm_fp8res(a, b, c512, M, 512, K)
    hb.gemm_fp8res(a, b384, c384, M, 384, K)
    torch.cuda.synchronize()
    wide = c512.view(M, 512)[:, :384].cpu().numpy()
    narrow = c384.view(M, 384).cpu().numpy()
    assert np.array_equal(wide.view(np.uint32), narrow.view(np.uint32))


def test_fp8res_padded_wide_form_equals_narrow_form():
    """N = 640 at M = 16384: round_up(N, 128) = 640 is not a multiple of 256,
    and on 256 CUs the 256 x 256 form (192 tiles, one round) beats the
    256 x 128 form (320 tiles, two rounds), so the wide kernel runs over B's
    planes padded to 768 columns, storing C's columns past N under
    predicates.  N = 384 runs the 256 x 128 form.  B's first 384 columns are
    whole scaling tiles in both, so the shared columns agree bit for bit, and
    C's columns past N (ldc = 704) keep their sentinel."""
    M, K, N, ldc = 16384, 256, 640, 704
    a = torch.empty(M * K, dtype=torch.float32, device="cuda")
    b = torch.empty(K * N, dtype=torch.float32, device="cuda")
    hb.fill_random(a, M * K, hb.GEN_NORMAL_OUTLIER, 21)
    hb.fill_random(b, K * N, hb.GEN_NORMAL_OUTLIER, 22)
    b384 = b.view(K, N)[:, :384].contiguous().view(-1)
    c640 = torch.full((M * ldc,), 7.5, dtype=torch.float32, device="cuda")
    c384 = torch.empty(M * 384, dtype=torch.float32, device="cuda")
    hb.gemm_fp8res(a, b, c640, M, N, K, ldc=ldc)
    hb.gemm_fp8res(a, b384, c384, M, 384, K)
    torch.cuda.synchronize()
    wide = c640.view(M, ldc).cpu().numpy()
    narrow = c384.view(M, 384).cpu().numpy()
    assert np.array_equal(wide[:, :384].view(np.uint32), narrow.view(np.uint32))
    assert np.all(wide[:, N:] == 7.5)
    # every column of the wide run within the algorithm's bound on a row band;
    # the band is one row of 128 x 128 scaling tiles, so quantising it alone
    # gives the scales the whole of A gets
    rows = slice(4096, 4224)
    A = a.view(M, K)[rows].cpu().numpy()
    B = b.view(K, N).cpu().numpy()
    Aq, Ar, _ = oracle.fp8_quantize(np.ascontiguousarray(A))
    BqT, BrT, _ = oracle.fp8_quantize(np.ascontiguousarray(B.T))
    Bq, Br = BqT.T.astype(np.float64), BrT.T.astype(np.float64)
    Aq, Ar = Aq.astype(np.float64), Ar.astype(np.float64)
    ideal = Aq @ Bq + Aq @ Br + Ar @ Bq
    absum = np.abs(Aq) @ np.abs(Bq) + np.abs(Aq) @ np.abs(Br) + np.abs(Ar) @ np.abs(Bq)
    err = np.abs(wide[rows, :N] - ideal)
    assert np.all(err <= gamma(3 * K + 2) * absum + fp8_block_term(Aq, Bq))


def test_fp8res_identity_b_reproduces_quantised_a():
    """B = I: every output is Aq + Ar of one A element (a single FP32 add of
    two exact terms), so C equals float32(Aq + Ar) bit for bit."""
    M, K = 256, 128
    a = torch.empty(M * K, device="cuda")
    hb.fill_random(a, M * K, hb.GEN_NORMAL_OUTLIER, 9)
    b = torch.eye(K, dtype=torch.float32, device="cuda").reshape(-1).contiguous()
    c = torch.empty(M * K, device="cuda")
    hb.gemm_fp8res(a, b, c, M, K, K)
    torch.cuda.synchronize()
    A = a.cpu().numpy().reshape(M, K)
    C = c.cpu().numpy().reshape(M, K)
    Aq, Ar, e = oracle.fp8_quantize(A)
    expect = (Aq.astype(np.float64) + Ar.astype(np.float64)).astype(np.float32)
    bad = C != expect
    scale = np.repeat(np.repeat(2.0 ** e, 128, 0), 128, 1)[:M, :K]
    sub = (Aq != 0) & (np.abs(Aq) < 2.0 ** -6 * scale)   # e4m3-subnormal q
    assert not bad.any(), (f"{bad.sum()} mismatches; {(bad & sub).sum()} at e4m3-subnormal q, "
                           f"{sub.sum()} subnormal q in total")


def test_fp8res_without_subnormals_matches_algorithm():
    """Entries with |x| in [0.5, 1) keep every x / 2^e in the e4m3 normal range."""
    M, N, K = 256, 128, 512
    g = torch.Generator(device="cuda").manual_seed(4)
    def draw(n):
        mag = 0.5 + 0.5 * torch.rand(n, device="cuda", generator=g)
        sgn = torch.where(torch.rand(n, device="cuda", generator=g) < 0.5, -1.0, 1.0)
        return (mag * sgn).float().contiguous()
    a = draw(M * K); b = draw(K * N)
    c = torch.empty(M * N, device="cuda")
    hb.gemm_fp8res(a, b, c, M, N, K)
    torch.cuda.synchronize()
    A = a.cpu().numpy().reshape(M, K); B = b.cpu().numpy().reshape(K, N)
    C = c.cpu().numpy().reshape(M, N)
    Aq, Ar, _ = oracle.fp8_quantize(A)
    BqT, BrT, _ = oracle.fp8_quantize(np.ascontiguousarray(B.T))
    Aq, Ar = Aq.astype(np.float64), Ar.astype(np.float64)
    Bq, Br = BqT.T.astype(np.float64), BrT.T.astype(np.float64)
    ideal = Aq @ Bq + Aq @ Br + Ar @ Bq
    absum = np.abs(Aq) @ np.abs(Bq) + np.abs(Aq) @ np.abs(Br) + np.abs(Ar) @ np.abs(Bq)
    assert np.all(np.abs(C - ideal) <= gamma(3 * K + 2) * absum)


def test_fp8res_full_size_outliers_error_report():
    """M = N = 16384, K = 4096, normal(0,1) with 1% x1e3 outliers: FP8+residual
    error against FP64 on a row band, beside BF16x3 and FP32 on the same inputs."""
    M = N = 16384
    K = 4096
    a = torch.empty(M * K, device="cuda"); b = torch.empty(K * N, device="cuda")
    hb.fill_random(a, M * K, hb.GEN_NORMAL_OUTLIER, 5)
    hb.fill_random(b, K * N, hb.GEN_NORMAL_OUTLIER, 6)
    c8 = torch.empty(M * N, device="cuda")
    ws = hb.gemm_fp8res(a, b, c8, M, N, K)
    del ws
    c3 = torch.empty(M * N, device="cuda")
    hb.gemm_bf16x3(a, b, c3, M, N, K)
    c32 = torch.empty(M * N, device="cuda")
    hb.sgemm_strided(a, b, c32, M, N, K, K, N, N)
    torch.cuda.synchronize()
    rows = dict(row_begin=8192, row_end=8192 + 256)
    e8 = hb.verify_f64(a, b, c8, M, N, K, **rows)
    e3 = hb.verify_f64(a, b, c3, M, N, K, **rows)
    e32 = hb.verify_f64(a, b, c32, M, N, K, **rows)
    print(f"\nC5 rows 8192..8447: e_cw fp8res={e8['e_cw']:.3e} bf16x3={e3['e_cw']:.3e} "
          f"fp32={e32['e_cw']:.3e}; e_F fp8res={e8['e_F']:.3e} bf16x3={e3['e_F']:.3e} "
          f"fp32={e32['e_F']:.3e}", flush=True)
    assert e8["nonfinite"] == 0
    # BF16-level: within the unit roundoff of bf16 inputs (2^-8) normwise
    assert e8["e_F"] <= 2.0 ** -8 and e8["e_nw"] <= 2.0 ** -8
    # componentwise at full size: rows 8192..8207 x columns 0..2047 of the
    # same launch within the algorithm's bound against the FP64 product, from
    # the oracle's quantisation of their whole 128 x 128 scaling tiles
    r0, nr, nc = 8192, 16, 2048
    At = a.view(M, K)[r0:r0 + 128].cpu().numpy()
    Bt = b.view(K, N)[:, :nc].cpu().numpy()
    Aq, Ar, _ = oracle.fp8_quantize(At)
    BqT, BrT, _ = oracle.fp8_quantize(np.ascontiguousarray(Bt.T))
    Aq, Ar = Aq[:nr].astype(np.float64), Ar[:nr].astype(np.float64)
    Bq, Br = BqT.T.astype(np.float64), BrT.T.astype(np.float64)
    A64, B64 = At[:nr].astype(np.float64), Bt.astype(np.float64)
    absum = np.abs(Aq) @ np.abs(Bq) + np.abs(Aq) @ np.abs(Br) + np.abs(Ar) @ np.abs(Bq)
    dA, dB = A64 - Aq, B64 - Bq
    algo = (2.0 ** -8 * (np.abs(Aq) @ np.abs(dB)) + 2.0 ** -8 * (np.abs(dA) @ np.abs(Bq))
            + np.abs(dA) @ np.abs(dB))
    bound = algo + gamma(3 * K + 2) * absum + fp8_block_term(Aq, Bq)
    Cs = c8.view(M, N)[r0:r0 + nr, :nc].cpu().numpy()
    R = A64 @ B64
    assert np.all(np.abs(Cs - R) <= bound)
    print(f"componentwise: max |C - R| / bound = {(np.abs(Cs - R) / bound).max():.3f}", flush=True)


def test_fp8res_non_finite_inputs_stay_in_their_row_and_column():
    """A NaN or an infinity in row i of A (column j of B) makes every output
    of row i (column j) of C non-finite and changes no other output: tile
    scales come from the finite entries only.  Every other output is within
    the bound of the algorithm applied to the inputs with the non-finite
    entries set to zero (the quantisation of the finite entries is the same),
    and both entry points agree bit for bit."""
    M = N = K = 384
    rng = np.random.default_rng(17)
    A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    A[5, 300] = np.nan; A[200, 17] = np.inf; A[130, 129] = -np.inf
    B[40, 7] = np.nan; B[100, 250] = np.inf
    bad_rows, bad_cols = [5, 130, 200], [7, 250]
    a = torch.from_numpy(A.ravel()).cuda(); b = torch.from_numpy(B.ravel()).cuda()
    c = torch.empty(M * N, dtype=torch.float32, device="cuda")
    c2 = torch.empty_like(c)
    hb.gemm_fp8res(a, b, c, M, N, K)
    hb.gemm_fp8res_noalloc(a, b, c2, M, N, K)
    torch.cuda.synchronize()
    C = c.cpu().numpy().reshape(M, N)
    assert np.array_equal(C, c2.cpu().numpy().reshape(M, N), equal_nan=True)
    fin = np.ones((M, N), bool)
    fin[bad_rows, :] = False
    fin[:, bad_cols] = False
    assert not np.isfinite(C[~fin]).any()
    assert np.isfinite(C[fin]).all()
    Az = np.where(np.isfinite(A), A, 0).astype(np.float32)
    Bz = np.where(np.isfinite(B), B, 0).astype(np.float32)
    Aq, Ar, _ = oracle.fp8_quantize(Az)
    BqT, BrT, _ = oracle.fp8_quantize(np.ascontiguousarray(Bz.T))
    Bq, Br = BqT.T.astype(np.float64), BrT.T.astype(np.float64)
    Aq, Ar = Aq.astype(np.float64), Ar.astype(np.float64)
    ideal = Aq @ Bq + Aq @ Br + Ar @ Bq
    absum = np.abs(Aq) @ np.abs(Bq) + np.abs(Aq) @ np.abs(Br) + np.abs(Ar) @ np.abs(Bq)
    bound = gamma(3 * K + 2) * absum + fp8_block_term(Aq, Bq)
    assert np.all(np.abs(C - ideal)[fin] <= bound[fin])
