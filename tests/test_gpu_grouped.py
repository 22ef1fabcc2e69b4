"""GPU parity of the grouped persistent GEMM (FP32-MFMA and BF16x3), via the C-ABI.

No reference function exists for grouping; it generalises the reference's
NITER slices in one launch (apps/sgemm/kernel.cpp:293-297).  Pinned by
 * exactness on small-integer inputs (the reference pattern's values 0..6 and
   0..2, so every partial sum is an exact integer), against an integer
   product computed in float64;
 * per-problem FP64 bounds: FP32 |C - R| <= gamma_{K+2} S; BF16x3
   |C - R| <= (3.1 * 2^-16 + gamma_{3K+2}) S (S = |A||B|);
 * determinism of split-K (fixed-order reduction) and of hipGraph replays.
"""
import os

import numpy as np
import pytest

import oracle

torch = pytest.importorskip("torch")
import hb_hammerbench_amd as hb  # noqa: E402

pytestmark = pytest.mark.gpu
U = 2.0 ** -24


def gamma(n):
    return n * U / (1 - n * U)


def make_group(shapes, kind="uniform", seed=5):
    """Allocates A, B, C per problem inside three flat buffers (16-byte aligned)."""
    offs_a, offs_b, offs_c = [], [], []
    ta = tb = tc = 0
    for (M, N, K) in shapes:
        offs_a.append(ta); offs_b.append(tb); offs_c.append(tc)
        ta += (M * K + 63) // 64 * 64
        tb += (K * N + 63) // 64 * 64
        tc += (M * N + 63) // 64 * 64
    a = torch.empty(ta, dtype=torch.float32, device="cuda")
    b = torch.empty(tb, dtype=torch.float32, device="cuda")
    c = torch.full((tc,), float("nan"), dtype=torch.float32, device="cuda")
    if kind == "pattern":
        hb.fill_pattern(a, b, min(ta, tb))
        if ta > tb:
            hb.fill_pattern(a[tb:], torch.empty(ta - tb, device="cuda"), ta - tb, first=tb)
        elif tb > ta:
            hb.fill_pattern(torch.empty(tb - ta, device="cuda"), b[ta:], tb - ta, first=ta)
    else:
        k = {"uniform": hb.GEN_UNIFORM, "loguniform": hb.GEN_LOGUNIFORM}[kind]
        hb.fill_random(a, ta, k, seed)
        hb.fill_random(b, tb, k, seed + 1)
    probs = []
    views = []
    for (M, N, K), oa, ob, oc in zip(shapes, offs_a, offs_b, offs_c):
        A = a[oa:oa + M * K].view(M, K); B = b[ob:ob + K * N].view(K, N)
        C = c[oc:oc + M * N].view(M, N)
        probs.append((A, B, C, M, N, K))
        views.append((A, B, C))
    return probs, views, (a, b, c)


# N and K are multiples of 4 (the grouped API's 16-byte load requirement)
RAGGED = [(64, 64, 64), (100, 36, 52), (128, 256, 64), (200, 68, 1000), (64, 2048, 128),
          (1, 4, 4), (333, 12, 4), (512, 128, 2048), (130, 132, 132)]


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
def test_grouped_integer_inputs_exact(prec):
    probs, views, _ = make_group(RAGGED, "pattern")
    plan = hb.GroupedPlan(probs, prec)
    plan.run()
    torch.cuda.synchronize()
    for (A, B, C), shape in zip(views, RAGGED):
        expect = (A.double() @ B.double()).float()
        assert torch.equal(C, expect), shape


@pytest.mark.parametrize("tile,split", [(0, 1), (1, 1), (2, 1), (3, 1), (0, 2)])
def test_grouped_bf16x3_identity_reproduces_oracle_split(tile, split):
    """BF16x3 with B = I (first problem) and with A = I (second): each output
    is lo + 0 + hi of the other operand's element, and hi + lo is exact in
    FP32, so C equals the oracle's split summed in float32 bit for bit.  Pins
    the in-register split of the A path and of the B path for every tile
    shape, with ragged M, N and K (132 is not a multiple of the 32-deep
    stage); with split-K the other splits add exact zeros."""
    M, K, N = 200, 132, 260

    def loguniform(n, seed):
        x = torch.empty(n, dtype=torch.float32, device="cuda")
        hb.fill_random(x, n, hb.GEN_LOGUNIFORM, seed)
        return x

    a1 = loguniform(M * K, 71).view(M, K)
    b2 = loguniform(K * N, 72).view(K, N)
    eye = torch.eye(K, dtype=torch.float32, device="cuda").contiguous()
    c1 = torch.full((M, K), float("nan"), device="cuda")
    c2 = torch.full((K, N), float("nan"), device="cuda")
    plan = hb.GroupedPlan([(a1, eye, c1, M, K, K), (eye, b2, c2, K, N, K)], hb.PREC_BF16X3,
                          tile=tile, split_k=split)
    plan.run()
    torch.cuda.synchronize()
    for x, c in ((a1, c1), (b2, c2)):
        xh = x.cpu().numpy()
        hi, lo = oracle.split_bf16(xh)
        want = oracle.bf16_bits_to_f32(hi) + oracle.bf16_bits_to_f32(lo)
        assert (want != xh).any()
        assert np.array_equal(c.cpu().numpy(), want), (tile, split)


FP32_ONLY_TILES = (hb.TILE_128x128_4x1, hb.TILE_128x128_BK32)


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
@pytest.mark.parametrize("tile,split", [(0, 1), (1, 1), (2, 1), (3, 1), (0, 3), (1, 4), (2, 2),
                                        (5, 1), (5, 3), (6, 1), (6, 2), (7, 1), (8, 1), (9, 1)])
def test_grouped_forced_configs_within_bound(prec, tile, split):
    """Every tile configuration (256-thread tiles with and without split-K,
    the 4 x 1 wave and BK 32 variants, the 8-wave 128 x 128, 256 x 128 and
    128 x 256 large tiles) on ragged problems: within the FP64 bound, and a
    second run identical to the first."""
    probs, views, _ = make_group(RAGGED, "uniform", seed=7 + tile)
    if prec == hb.PREC_BF16X3 and tile in FP32_ONLY_TILES:
        with pytest.raises(hb.HbsError) as e:
            hb.GroupedPlan(probs, prec, tile=tile, split_k=split)
        assert e.value.code == -1
        return
    plan = hb.GroupedPlan(probs, prec, tile=tile, split_k=split)
    info = plan.info()
    if split > 1:
        assert info["split_tiles"] > 0
    plan.run()
    torch.cuda.synchronize()
    first = [C.clone() for (_, _, C) in views]
    plan.run()  # split-K counters reset themselves: a second run is identical
    torch.cuda.synchronize()
    for (A, B, C), c0, (M, N, K) in zip(views, first, RAGGED):
        assert torch.equal(C, c0), (M, N, K)
        R, S = oracle.gemm_f64(A.cpu().numpy(), B.cpu().numpy())
        if prec == hb.PREC_FP32:
            bound = gamma(K + 2) * S
        else:
            bound = (3.1 * 2.0 ** -16 + gamma(3 * K + 2)) * S
        assert np.all(np.abs(C.cpu().numpy() - R) <= bound), (M, N, K)


def test_grouped_plan_info_counts_tiles():
    shapes = [(200, 68, 64), (64, 64, 64), (129, 260, 32)]
    probs, _, _ = make_group(shapes)
    plan = hb.GroupedPlan(probs, hb.PREC_FP32, tile=0, split_k=1)
    info = plan.info()
    expect = sum(((M + 127) // 128) * ((N + 127) // 128) for M, N, _ in shapes)
    assert info["tiles"] == expect and info["work_items"] == expect
    assert info["problems"] == 3 and info["workgroups"] >= 1


# Shapes for the large-tile path: whole problems in 256 x 256 tiles, ragged at
# the bottom and right (clamped loads) or not, plus one K that is not a
# multiple of 32, which runs on a second large-tile launch that zero-fills
# past K.
BIG = [(300, 260, 256), (512, 512, 320), (257, 1024, 256), (256, 256, 512), (1000, 1028, 768),
       (1024, 300, 1024), (512, 512, 100), (64, 2048, 256)]


def big_tiles(shapes):
    """256 x 256 tiles of these problems on the large-tile launches."""
    return sum(((M + 255) // 256) * ((N + 255) // 256) for M, N, _ in shapes)


# Tuning lines that send the thin, long-K class (M, N about 64, K about 1024
# or 2048) to 64 x 64 tiles with split-K 4 on the 256-thread launch, so that a
# group also exercises that launch.
THIN_SPLIT = "".join(f"{p} 64 64 {k} 1 4 0\n" for p in ("fp32", "bf16x3") for k in (1024, 2048))


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
def test_grouped_large_tiles_integer_exact(prec):
    probs, views, _ = make_group(BIG, "pattern")
    plan = hb.GroupedPlan(probs, prec, tile=hb.TILE_256x256, split_k=1)
    plan.run()
    torch.cuda.synchronize()
    for (A, B, C), shape in zip(views, BIG):
        assert torch.equal(C, (A.double() @ B.double()).float()), shape


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
@pytest.mark.parametrize("mode", ["forced", "forced-split2", "default"])
def test_grouped_large_tiles_within_bound(prec, mode):
    """Whole problems on the large-tile launch (edge tiles with idle waves,
    per-XCD queues) and K = 100 on the zero-filling large-tile launch: every
    problem within its FP64 bound, a second run identical (the queues reset
    themselves), and the plan's tiles all 256 x 256 (split-K does not apply)."""
    probs, views, _ = make_group(BIG, "uniform", seed=29)
    if mode == "default":
        plan = hb.GroupedPlan(probs, prec)
    else:
        plan = hb.GroupedPlan(probs, prec, tile=hb.TILE_256x256, split_k=2 if "split" in mode else 1)
    info = plan.info()
    nbig = big_tiles(BIG)
    assert info["tiles"] == nbig and info["work_items"] == nbig and info["split_tiles"] == 0
    plan.run()
    torch.cuda.synchronize()
    first = [C.clone() for (_, _, C) in views]
    plan.run()
    torch.cuda.synchronize()
    for (A, B, C), c0, (M, N, K) in zip(views, first, BIG):
        assert torch.equal(C, c0), (M, N, K)
        R, S = oracle.gemm_f64(A.cpu().numpy(), B.cpu().numpy())
        if prec == hb.PREC_FP32:
            bound = gamma(K + 2) * S
        else:
            bound = (3.1 * 2.0 ** -16 + gamma(3 * K + 2)) * S
        assert np.all(np.abs(C.cpu().numpy() - R) <= bound), (M, N, K)


def test_grouped_large_tiles_identity_reproduces_oracle_split():
    """BF16x3 through the large-tile launch with ragged edges: with B = I the
    output is the oracle's hi + lo of A bit for bit (rows 256..299 come from a
    tile that clamps its loads), and with A = I that of B (columns 256..259
    likewise)."""
    M, K, N = 300, 512, 260

    def loguniform(n, seed):
        x = torch.empty(n, dtype=torch.float32, device="cuda")
        hb.fill_random(x, n, hb.GEN_LOGUNIFORM, seed)
        return x

    a1 = loguniform(M * K, 81).view(M, K)
    b2 = loguniform(K * N, 82).view(K, N)
    eye = torch.eye(K, dtype=torch.float32, device="cuda").contiguous()
    c1 = torch.full((M, K), float("nan"), device="cuda")
    c2 = torch.full((K, N), float("nan"), device="cuda")
    plan = hb.GroupedPlan([(a1, eye, c1, M, K, K), (eye, b2, c2, K, N, K)], hb.PREC_BF16X3,
                          tile=hb.TILE_256x256, split_k=1)
    assert plan.info()["tiles"] == 2 * 2 + 2 * 2  # both problems whole, in 256 x 256 tiles
    plan.run()
    torch.cuda.synchronize()
    for x, c in ((a1, c1), (b2, c2)):
        xh = x.cpu().numpy()
        hi, lo = oracle.split_bf16(xh)
        want = oracle.bf16_bits_to_f32(hi) + oracle.bf16_bits_to_f32(lo)
        assert np.array_equal(c.cpu().numpy(), want)


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
def test_grouped_graph_replays_match_eager(prec, tmp_path):
    # Every launch in one graph: the large-tile launch; (64, 64, 2048) on the
    # 256-thread launch with split-K (tuning file); K = 100 on the
    # zero-filling large-tile launch
    shapes = [(256, 192, 512), (64, 64, 2048), (96, 320, 64), (1024, 64, 1024), (520, 772, 512),
              (200, 300, 100)]
    probs, views, bufs = make_group(shapes, "uniform", seed=3)
    path = tmp_path / "tune.txt"
    path.write_text(THIN_SPLIT)
    plan = hb.GroupedPlan(probs, prec, tuning_file=str(path))
    info = plan.info()
    big = [s for i, s in enumerate(shapes) if i != 1]
    assert info["tiles"] == big_tiles(big) + 1
    assert info["split_tiles"] == 1 and info["work_items"] == info["tiles"] + 3
    plan.run()
    torch.cuda.synchronize()
    eager = [C.clone() for (_, _, C) in views]
    g = plan.graph()
    for rep in range(3):
        bufs[2].zero_()             # eager work between replays
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for (_, _, C), e in zip(views, eager):
            assert torch.equal(C, e), rep
    g.close()


def test_grouped_tuning_file_is_applied(tmp_path):
    shapes = [(64, 64, 2048)] * 3
    probs, views, _ = make_group(shapes, "uniform", seed=11)
    path = tmp_path / "tune.txt"
    path.write_text("# test\nfp32 64 64 2048 1 4 1000.0\n")
    plan = hb.GroupedPlan(probs, hb.PREC_FP32, tuning_file=str(path))
    info = plan.info()
    assert info["split_tiles"] == 3 and info["work_items"] == 12
    plan.run()
    torch.cuda.synchronize()
    for (A, B, C) in views:
        R, S = oracle.gemm_f64(A.cpu().numpy(), B.cpu().numpy())
        assert np.all(np.abs(C.cpu().numpy() - R) <= gamma(2048 + 2) * S)
    with pytest.raises(hb.HbsError):
        hb.GroupedPlan(probs, hb.PREC_FP32, tuning_file=str(tmp_path / "missing.txt"))
    # a large-tile entry (tile id 4) moves its class to the 512-thread launch
    shapes = [(512, 512, 512), (64, 64, 64)]
    probs, views, _ = make_group(shapes, "uniform", seed=12)
    path.write_text("fp32 512 512 512 4 1 0\n")
    plan = hb.GroupedPlan(probs, hb.PREC_FP32, tuning_file=str(path))
    info = plan.info()
    assert info["tiles"] == 4 + 1 and info["work_items"] == 4 + 1
    plan.run()
    torch.cuda.synchronize()
    for (A, B, C), (M, N, K) in zip(views, shapes):
        R, S = oracle.gemm_f64(A.cpu().numpy(), B.cpu().numpy())
        assert np.all(np.abs(C.cpu().numpy() - R) <= gamma(K + 2) * S)


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
def test_grouped_512_problem_workload(prec):
    """The 512-problem group, M, N, K uniform over {64, ..., 2048}: every
    problem's componentwise error within its bound (FP64 on the device)."""
    rng = np.random.default_rng(2024)
    shapes = [tuple(int(x) for x in rng.integers(1, 33, 3) * 64) for _ in range(512)]
    probs, views, _ = make_group(shapes, "uniform", seed=17)
    plan = hb.GroupedPlan(probs, prec)
    g = plan.graph()
    g.launch()
    torch.cuda.synchronize()
    worst = 0.0
    for (A, B, C), (M, N, K) in zip(views, shapes):
        st = hb.verify_f64(A, B, C, M, N, K)
        lim = gamma(K + 2) if prec == hb.PREC_FP32 else 3.1 * 2.0 ** -16 + gamma(3 * K + 2)
        assert st["nonfinite"] == 0 and st["e_cw"] <= lim, (M, N, K, st["e_cw"])
        worst = max(worst, st["e_cw"])
    print(f"\n512-problem group prec={prec}: worst e_cw {worst:.3e}", flush=True)
    g.close()


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
def test_grouped_many_small_problems_through_the_queues(prec, tmp_path):
    """300 problems of random shapes, a quarter with K not a multiple of 32
    (the zero-filling large-tile launch), some thin and long in K on the
    256-thread launch with split-K (a tuning file): every launch runs, with
    thousands of queue fetches per run.
    Two graph replays must agree bit for bit with an eager run (the queues and
    split counters reset themselves) and every problem is within its FP64
    bound."""
    rng = np.random.default_rng(2024)
    shapes = []
    for i in range(300):
        M = int(rng.integers(1, 200))
        N = 4 * int(rng.integers(1, 50))
        K = 4 * int(rng.integers(1, 60)) if i % 4 == 0 else 32 * int(rng.integers(1, 9))
        if i % 50 == 0:
            M, N, K = 64, 64, 1028  # thin and long: split-K on the 256-thread launch
        shapes.append((M, N, K))
    probs, views, _ = make_group(shapes, "uniform", seed=71)
    path = tmp_path / "tune.txt"
    path.write_text(THIN_SPLIT)
    plan = hb.GroupedPlan(probs, prec, tuning_file=str(path))
    info = plan.info()
    assert info["split_tiles"] > 0 and info["work_items"] > info["tiles"]
    plan.run()
    torch.cuda.synchronize()
    eager = [C.clone() for (_, _, C) in views]
    g = plan.graph()
    for rep in range(2):
        for (_, _, C) in views:
            C.fill_(float("nan"))
        g.launch()
        torch.cuda.synchronize()
        for (_, _, C), e, shape in zip(views, eager, shapes):
            assert torch.equal(C, e), (rep, shape)
    g.close()
    for (A, B, C), (M, N, K) in zip(views, shapes):
        R, S = oracle.gemm_f64(A.cpu().numpy(), B.cpu().numpy())
        if prec == hb.PREC_FP32:
            bound = gamma(K + 2) * S
        else:
            bound = (3.1 * 2.0 ** -16 + gamma(3 * K + 2)) * S
        assert np.all(np.abs(C.cpu().numpy() - R) <= bound), (M, N, K)


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
def test_grouped_split_64_long_k(prec):
    """The long-K route for a product with few output tiles: one problem,
    forced 128 x 128 tiles with split-K 64 (4 tiles x 64 K ranges).  Within
    the FP64 bound and identical between an eager run and a graph replay."""
    shapes = [(256, 256, 16384)]
    probs, views, _ = make_group(shapes, "uniform", seed=81)
    plan = hb.GroupedPlan(probs, prec, tile=hb.TILE_128x128, split_k=64)
    info = plan.info()
    assert info["tiles"] == 4 and info["work_items"] == 4 * 64 and info["split_tiles"] == 4
    plan.run()
    torch.cuda.synchronize()
    (A, B, C), = views
    eager = C.clone()
    C.fill_(float("nan"))
    g = plan.graph()
    g.launch()
    torch.cuda.synchronize()
    assert torch.equal(C, eager)
    g.close()
    M, N, K = shapes[0]
    R, S = oracle.gemm_f64(A.cpu().numpy(), B.cpu().numpy())
    if prec == hb.PREC_FP32:
        bound = gamma(K + 2) * S
    else:
        bound = (3.1 * 2.0 ** -16 + gamma(3 * K + 2)) * S
    assert np.all(np.abs(C.cpu().numpy() - R) <= bound)


@pytest.mark.parametrize("prec,ks", [(hb.PREC_FP32, (32, 64, 96, 128, 256)),
                                     (hb.PREC_BF16X3, (32, 64, 96, 128, 256)),
                                     (hb.PREC_FP32, (4, 36, 68, 100, 260)),
                                     (hb.PREC_BF16X3, (4, 36, 68, 100, 260))])
def test_grouped_chained_tiles_match_problems_run_alone(prec, ks):
    """The large-tile launch runs a workgroup's tiles back to back, loading the
    next tile's first stages during the current tile's last steps.  A tile's
    arithmetic does not depend on which tile ran before it, so a group of
    about a thousand 256 x 256 tiles with K of 1, 2, 3, 4 and 8 or 9 stages
    and ragged M and N must give, bit for bit, what each problem gives in a
    plan of its own (one or a few tiles per workgroup).  The second K set (not
    multiples of 32) runs on the zero-filling large-tile launch."""
    rng = np.random.default_rng(77)
    shapes = []
    for i in range(130):
        M = int(rng.integers(1, 1100))
        N = 4 * int(rng.integers(1, 275))
        K = ks[i % 5]
        shapes.append((M, N, K))
    probs, views, _ = make_group(shapes, "uniform", seed=91)
    plan = hb.GroupedPlan(probs, prec)
    info = plan.info()
    tiles = sum(-(-M // 256) * -(-N // 256) for M, N, K in shapes)
    assert info["tiles"] == tiles and tiles > 3 * 256 and info["split_tiles"] == 0
    plan.run()
    torch.cuda.synchronize()
    together = [C.clone() for (_, _, C) in views]
    for pr, (_, _, C), t, shape in zip(probs, views, together, shapes):
        C.fill_(float("nan"))
        alone = hb.GroupedPlan([pr], prec)
        alone.run()
        torch.cuda.synchronize()
        assert torch.equal(C, t), shape
    for (A, B, C), (M, N, K) in zip(views[:10], shapes[:10]):
        R, S = oracle.gemm_f64(A.cpu().numpy(), B.cpu().numpy())
        if prec == hb.PREC_FP32:
            bound = gamma(K + 2) * S
        else:
            bound = (3.1 * 2.0 ** -16 + gamma(3 * K + 2)) * S
        assert np.all(np.abs(C.cpu().numpy() - R) <= bound), (M, N, K)


def test_plan_outlives_its_graphs():
    """A plan with a live graph refuses to be destroyed (HBS_ERR_IN_USE, its
    device buffers stay valid and the graph still replays correctly); once the
    graph is destroyed the plan can be."""
    import ctypes
    shapes = [(128, 96, 64), (64, 256, 96)]
    probs, views, keep = make_group(shapes, "pattern")
    plan = hb.GroupedPlan(probs, hb.PREC_FP32)
    g = plan.graph()
    L = hb.lib()
    assert L.hbs_grouped_plan_destroy(plan.handle) == -8
    g.launch()
    torch.cuda.synchronize()
    for (A, B, C), (M, N, K) in zip(views, shapes):
        R = A.double().cpu().numpy() @ B.double().cpu().numpy()
        assert np.array_equal(C.cpu().numpy(), R.astype(np.float32))
    h2 = ctypes.c_void_p()
    assert L.hbs_grouped_graph_create(plan.handle, ctypes.byref(h2)) == 0
    g.close()
    assert L.hbs_grouped_plan_destroy(plan.handle) == -8  # the second graph is live
    assert L.hbs_graph_destroy(h2) == 0
    plan.close()
    assert plan.handle is None


def test_plan_and_graph_cycles_do_not_leak_device_memory():
    """40 cycles of plan create, graph capture, launch, graph and plan destroy
    on a group whose plan holds about 0.5 GB of device memory (2000 small
    problems as 3999 128 x 128 tiles split in two: their FP32 partials):
    free device memory returns to where it was, within 8 MiB, which one
    leaked plan would exceed sixty times over.  The first cycle runs before
    the baseline, so the library's one-time setup is not counted."""
    shapes = [(64 + 64 * (i % 4), 64 + 64 * (i % 3), 128) for i in range(2000)]
    probs, views, keep = make_group(shapes, "uniform")

    def cycle():
        plan = hb.GroupedPlan(probs, hb.PREC_BF16X3, tile=0, split_k=2)
        g = plan.graph()
        g.launch()
        torch.cuda.synchronize()
        g.close()
        plan.close()

    cycle()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    plan = hb.GroupedPlan(probs, hb.PREC_BF16X3, tile=0, split_k=2)
    per_plan = plan.info()
    plan.close()
    for _ in range(40):
        cycle()
    free1 = torch.cuda.mem_get_info()[0]
    assert free0 - free1 < 8 << 20, (free0 - free1, per_plan)


EDGE_REMS = (40, 64, 100, 128, 148, 192, 228, 256)


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
@pytest.mark.parametrize("K", [64, 100])
def test_edge_tile_arrangements_exact(prec, K):
    """Every valid region an edge tile of the large-tile launch can have: M and
    N remainders falling into each count of 64-row and 64-column units (1..4),
    so that every wave arrangement of the edge tiles -- pieces of 64 x 128 and
    64 x 64 spread over the SIMDs, or none -- computes and stores its part.
    Small-integer inputs: the result is the exact integer product."""
    shapes = []
    for i, rm in enumerate(EDGE_REMS):
        for j, rn in enumerate(EDGE_REMS):
            shapes.append((256 * ((i + j) % 2) + rm, 256 * (j % 2) + rn, K))
    probs, views, keep = make_group(shapes, "pattern")
    plan = hb.GroupedPlan(probs, prec)
    plan.run()
    torch.cuda.synchronize()
    for (A, B, C), (M, N, KK) in zip(views, shapes):
        R = A.double() @ B.double()
        assert torch.equal(C, R.float()), (M, N, KK)


@pytest.mark.parametrize("N", [3900, 4000, 3830])
def test_square_edge_tiles_exact(N):
    """Square kernels take the same edge arrangement on their last tile row and
    column when they run 256 x 256 tiles: remainders of 60 (K not a multiple
    of 32: zero-filling loads), 160 (clamped loads) and 246 (N not a multiple
    of 4: predicated loads)."""
    n = N * N
    a = torch.empty(n, dtype=torch.float32, device="cuda"); b = torch.empty_like(a)
    c = torch.empty_like(a)
    hb.fill_pattern(a, b, n)
    expect = oracle.pattern_slice(N, 0)
    for fn in (hb.sgemm, hb.sgemm_bf16x3):
        c.fill_(float("nan"))
        fn(a, b, c, N, 1)
        torch.cuda.synchronize()
        assert np.array_equal(c.view(N, N).cpu().numpy(), expect), (fn.__name__, N)


# (M, N, K, lda, ldb, ldc, float offset of A, of B): operands that allow no
# 16-byte loads (N, K, lda or ldb not a multiple of 4, or A / B one float off
# a 16-byte boundary) next to ones that do, in one plan
ODD_PROBLEMS = [
    (300, 301, 257, 257, 301, 301, 0, 0),
    (129, 67, 130, 131, 70, 69, 1, 3),
    (256, 256, 256, 256, 256, 256, 1, 0),       # aligned shape, A misaligned
    (64, 128, 96, 96, 128, 128, 0, 2),          # aligned shape, B misaligned
    (1, 1, 1, 1, 1, 1, 0, 0),
    (513, 260, 64, 64, 260, 260, 0, 0),         # aligned: the clamped launch
    (70, 90, 100, 100, 92, 90, 0, 0),           # aligned: zero fill past K
]


def odd_group(kind, seed=9):
    offs, ta, tb, tc = [], 0, 0, 0
    for (M, N, K, lda, ldb, ldc, oa, ob) in ODD_PROBLEMS:
        offs.append((ta + oa, tb + ob, tc))
        ta += (M * lda + oa + 63) // 64 * 64
        tb += (K * ldb + ob + 63) // 64 * 64
        tc += (M * ldc + 63) // 64 * 64
    a = torch.empty(ta, dtype=torch.float32, device="cuda")
    b = torch.empty(tb, dtype=torch.float32, device="cuda")
    c = torch.full((tc,), -3.0, dtype=torch.float32, device="cuda")
    if kind == "pattern":
        hb.fill_pattern(a, torch.empty(ta, device="cuda"), ta)
        hb.fill_pattern(torch.empty(tb, device="cuda"), b, tb)
    else:
        hb.fill_random(a, ta, hb.GEN_UNIFORM, seed)
        hb.fill_random(b, tb, hb.GEN_UNIFORM, seed + 1)
    probs, views = [], []
    for (M, N, K, lda, ldb, ldc, _, _), (oa, ob, oc) in zip(ODD_PROBLEMS, offs):
        A = a[oa:oa + M * lda].view(M, lda)
        B = b[ob:ob + K * ldb].view(K, ldb)
        C = c[oc:oc + M * ldc].view(M, ldc)
        probs.append((A, B, C, M, N, K, lda, ldb, ldc))
        views.append((A[:, :K], B[:, :N], C, M, N, K))
    return probs, views, (a, b, c)


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
def test_grouped_any_shape_and_alignment(prec):
    """Problems of any shape, leading dimensions and operand alignment in one
    plan (three large-tile launches: clamped, zero fill past K, predicated):
    exact on small-integer inputs, within the FP64 bounds on uniform ones, the
    columns between N and ldc untouched, and a graph replay equal to the eager
    run."""
    probs, views, keep = odd_group("pattern")
    plan = hb.GroupedPlan(probs, prec)
    plan.run()
    torch.cuda.synchronize()
    for (A, B, C, M, N, K) in views:
        R = A.double() @ B.double()
        assert torch.equal(C[:, :N], R.float()), (M, N, K)
        assert bool((C[:, N:] == -3.0).all()), (M, N, K)
    probs, views, keep = odd_group("uniform")
    plan = hb.GroupedPlan(probs, prec)
    plan.run()
    torch.cuda.synchronize()
    eager = keep[2].clone()
    for (A, B, C, M, N, K) in views:
        R, S = oracle.gemm_f64(A.cpu().numpy(), B.cpu().numpy())
        Cn = C[:, :N].cpu().numpy()
        bound = (gamma(K + 2) if prec == hb.PREC_FP32
                 else 3.1 * 2.0 ** -16 + gamma(3 * K + 2)) * S
        assert np.all(np.abs(Cn - R) <= bound), (M, N, K)
    g = plan.graph()
    for _ in range(2):
        keep[2].fill_(-3.0)
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(keep[2], eager)
    g.close()


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
@pytest.mark.parametrize("tile", [hb.TILE_128x128_8W, hb.TILE_256x128, hb.TILE_128x256])
def test_large_tile_shapes_chained_match_alone(prec, tile):
    """The 8-wave 128 x 128, 256 x 128 and 128 x 256 large tiles chain their
    tiles through one pipeline like the 256 x 256 one: a group with K of 1, 3,
    4 and 8 stages and K not a multiple of 32 (the zero-filling launch) gives,
    bit for bit, what each problem gives alone; the pattern is exact."""
    rng = np.random.default_rng(5 + tile)
    ks = (32, 96, 128, 256, 100, 260)
    shapes = [(int(rng.integers(1, 700)), 4 * int(rng.integers(1, 180)), ks[i % 6])
              for i in range(60)]
    probs, views, _ = make_group(shapes, "uniform", seed=13)
    plan = hb.GroupedPlan(probs, prec, tile=tile)
    plan.run()
    torch.cuda.synchronize()
    together = [C.clone() for (_, _, C) in views]
    for pr, (_, _, C), t, shape in zip(probs, views, together, shapes):
        C.fill_(float("nan"))
        hb.GroupedPlan([pr], prec, tile=tile).run()
        torch.cuda.synchronize()
        assert torch.equal(C, t), shape
    probs, views, _ = make_group(shapes[:20], "pattern")
    hb.GroupedPlan(probs, prec, tile=tile).run()
    torch.cuda.synchronize()
    for (A, B, C), shape in zip(views, shapes[:20]):
        assert torch.equal(C, (A.double() @ B.double()).float()), shape
