"""The grouped planner on the host (no GPU): hbs_grouped_plan_describe runs the
same planning as hbs_grouped_plan_create -- default, tuning file or forced
tile / split-K, work-item expansion, launch sizes -- without a device.
Expected counts follow from the planning rules:

  * by default and with tile 4 a problem takes a large-tile launch:
    ceil(M/256) ceil(N/256) tiles, no split-K; problems whose K is a multiple
    of 32 form one launch, the others a second one (zero fill past K);
  * with a forced tile 0..3 a problem takes the 256-thread launch with that
    tile and split-K s: a K range is ceil(K / s) rounded up to the stage depth
    (16 for FP32, 32 for BF16x3), giving ceil(K / range) work items per tile;
  * a problem whose operands allow no 16-byte loads (here: N or K not a
    multiple of 4, with lda = K and ldb = N) takes a third large-tile launch
    with per-element predicated loads, whatever the tuning file or forced tile;
  * workgroups: one per CU on each large-tile launch, two per CU on the other,
    never more than the launch's work items.
"""
import math

import numpy as np
import pytest

import hb_hammerbench_amd as hb

# tile id -> (BM, BN, large-tile slot or None, K stage FP32, K stage BF16x3 or None)
TILE = {0: (128, 128, None, 16, 32), 1: (64, 64, None, 16, 32), 2: (128, 64, None, 16, 32),
        3: (64, 128, None, 16, 32), 4: (256, 256, 0, 32, 32), 5: (128, 128, None, 16, None),
        6: (128, 128, None, 32, None), 7: (128, 128, 1, 32, 32), 8: (256, 128, 2, 32, 32),
        9: (128, 256, 3, 32, 32)}
EDGE_LAUNCH = 8  # the 256 x 256 launch with per-element predicated loads


def tiles_for(prec):
    return [t for t, cfg in TILE.items() if cfg[3 if prec == hb.PREC_FP32 else 4]]


def expected(shapes, prec, cus=256, tile=None, split=1):
    big_tiles = [0] * 9  # large-tile launches: 2 slot + (K not a multiple of 32), predicated
    small_tiles = small_items = split_tiles = 0
    for M, N, K in shapes:
        vec = N % 4 == 0 and K % 4 == 0
        t = 4 if tile is None or not vec else tile
        BM, BN, slot, st32, st3 = TILE[t]
        if slot is not None:
            g = EDGE_LAUNCH if not vec else 2 * slot + (0 if K % 32 == 0 else 1)
            big_tiles[g] += math.ceil(M / BM) * math.ceil(N / BN)
            continue
        stage = st32 if prec == hb.PREC_FP32 else st3
        n = math.ceil(M / BM) * math.ceil(N / BN)
        chunk = math.ceil(math.ceil(K / split) / stage) * stage
        ns = math.ceil(K / chunk)
        small_tiles += n
        small_items += n * ns
        split_tiles += n if ns > 1 else 0
    return {"tiles": sum(big_tiles) + small_tiles, "work_items": sum(big_tiles) + small_items,
            "split_tiles": split_tiles,
            "workgroups": sum(min(cus, b) for b in big_tiles) + min(2 * cus, small_items)}


MIXED = [(1000, 1028, 768), (64, 2048, 256), (300, 260, 256), (200, 68, 100),
         (64, 64, 2052), (512, 512, 100), (130, 132, 132), (1, 4, 4), (64, 520, 4000)]


def check(info, want, count):
    for k, v in want.items():
        assert info[k] == v, (k, info[k], v)
    assert info["problems"] == count and info["est_imbalance"] >= 1.0


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
@pytest.mark.parametrize("cus", [256, 8])
def test_describe_default_plan(prec, cus):
    info = hb.describe_plan(MIXED, prec, cus=cus)
    want = expected(MIXED, prec, cus)
    # every problem on the large-tile launches, no split-K
    assert want["tiles"] > 20 and want["split_tiles"] == 0
    assert want["workgroups"] == min(cus, 35) + min(cus, 8)  # tiles with K % 32 == 0, others
    check(info, want, len(MIXED))


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
@pytest.mark.parametrize("tile,split", [(0, 1), (1, 3), (2, 2), (3, 5), (4, 1), (4, 2), (5, 1),
                                        (5, 4), (6, 1), (6, 3), (7, 1), (8, 1), (9, 2)])
def test_describe_forced_tiles(prec, tile, split):
    if tile not in tiles_for(prec):
        with pytest.raises(hb.HbsError) as e:
            hb.describe_plan(MIXED, prec, tile=tile, split_k=split)
        assert e.value.code == -1  # the 4 x 1 and BK 32 tiles are FP32 only
        return
    info = hb.describe_plan(MIXED, prec, tile=tile, split_k=split)
    check(info, expected(MIXED, prec, tile=tile, split=split), len(MIXED))


def test_describe_tuning_file_with_every_tile_id(tmp_path):
    """A tuning file may name any tile configuration per size class; lines
    whose tile the precision does not have are ignored."""
    shapes = [(64, 64, 2048), (512, 512, 512), (300, 1000, 96), (2048, 128, 1024),
              (100, 1800, 64), (130, 260, 1000)]
    lines = ["fp32 64 64 2048 5 4 0", "fp32 512 512 512 6 2 0", "fp32 256 1024 128 7 1 0",
             "fp32 2048 128 1024 8 1 0", "fp32 128 2048 64 9 1 0",
             "bf16x3 64 64 2048 5 2 0", "bf16x3 512 512 512 8 1 0", "bf16x3 128 256 1024 7 1 0"]
    path = tmp_path / "t.txt"
    path.write_text("\n".join(lines) + "\n")
    fp32 = hb.describe_plan(shapes, hb.PREC_FP32, tuning_file=str(path))
    # tiles per shape: 128 x 128 (4 x 1) split 4; 128 x 128 (BK 32) split 2 on
    # 4 x 4 tiles; 8 waves of 128 x 128 on 3 x 8; 256 x 128 on 8 x 1; 128 x 256
    # on 1 x 8; the class of 130 x 260 x 1000 (128 256 1024) has no fp32 line:
    # 256 x 256 on 1 x 2
    assert fp32["tiles"] == 1 + 16 + 24 + 8 + 8 + 2
    assert fp32["split_tiles"] == 1 + 16
    assert fp32["work_items"] == 4 + 16 * 2 + 24 + 8 + 8 + 2
    b3 = hb.describe_plan(shapes, hb.PREC_BF16X3, tuning_file=str(path))
    # 64 x 64 x 2048 keeps the default (tile 5 is FP32 only); 512^3 on 256 x 128
    # (2 x 4); 130 x 260 x 1000 on 8 waves of 128 x 128 (2 x 3); the others on
    # 256 x 256 tiles
    assert b3["split_tiles"] == 0
    assert b3["tiles"] == 1 + 8 + 2 * 4 + 8 * 1 + 1 * 8 + 6


def test_describe_tuning_file_overrides_the_default(tmp_path):
    shapes = [(64, 64, 2048)] * 3 + [(512, 512, 512)]
    assert hb.describe_plan(shapes)["tiles"] == 3 + 4  # all on 256 x 256 tiles
    path = tmp_path / "tune.txt"
    path.write_text("# test\nfp32 64 64 2048 1 4 1000.0\n")
    info = hb.describe_plan(shapes, tuning_file=str(path))
    assert info["tiles"] == 3 + 4 and info["work_items"] == 3 * 4 + 4
    assert info["split_tiles"] == 3 and info["workgroups"] == 4 + 12
    # a class the table does not name keeps the default; BF16x3 ignores fp32 lines
    info = hb.describe_plan(shapes, hb.PREC_BF16X3, tuning_file=str(path))
    assert info["work_items"] == 3 + 4 and info["split_tiles"] == 0


def test_describe_rejects_bad_arguments(tmp_path):
    def code(*a, **k):
        with pytest.raises(hb.HbsError) as e:
            hb.describe_plan(*a, **k)
        return e.value.code
    assert code([]) == -1                                   # no problems
    assert code([(0, 64, 64)]) == -1
    assert code([(64, 64, 64)], precision=7) == -1
    assert code([(64, 64, 64)], tile=10) == -1
    assert code([(64, 64, 64)], precision=hb.PREC_BF16X3, tile=6) == -1
    assert code([(64, 64, 64)], tile=0, split_k=0) == -1
    assert code([(64, 64, 64)], tile=0, tuning_file=str(tmp_path / "t.txt")) == -1
    assert code([(64, 64, 64)], tuning_file=str(tmp_path / "missing.txt")) == -1
    assert code([(64, 64, 64)], cus=0) == -1


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
def test_describe_any_shape_takes_the_predicated_launch(prec):
    """N or K not a multiple of 4 is accepted: such problems run whole on the
    third large-tile launch (per-element predicated loads), also when a tile
    shape and split-K are forced for the others."""
    shapes = [(64, 62, 64), (300, 301, 257), (1, 1, 1), (512, 512, 512), (256, 255, 64)]
    for tile, split in ((None, 1), (0, 4), (4, 1)):
        info = hb.describe_plan(shapes, prec, tile=tile, split_k=split)
        check(info, expected(shapes, prec, tile=tile, split=split), len(shapes))
    info = hb.describe_plan(shapes[:2], prec, tile=1, split_k=8)
    assert info["split_tiles"] == 0 and info["tiles"] == 1 + 2 * 2


@pytest.mark.parametrize("seed", range(6))
def test_describe_random_groups(seed):
    """Random groups (K any multiple of 4, so both large-tile launches of
    BF16x3 and the 256-thread launch of FP32 all appear), random CU counts,
    default and forced plans: the dry run's counts follow the rules above."""
    rng = np.random.default_rng(seed)
    shapes = [(int(rng.integers(1, 3000)), 4 * int(rng.integers(1, 700)),
               4 * int(rng.integers(1, 1100))) for _ in range(int(rng.integers(1, 60)))]
    # a few shapes whose N or K is not a multiple of 4
    shapes += [(int(rng.integers(1, 600)), int(rng.integers(1, 700)), int(rng.integers(1, 900)))
               for _ in range(int(rng.integers(0, 6)))]
    cus = int(rng.choice([8, 24, 104, 256]))
    for prec in (hb.PREC_FP32, hb.PREC_BF16X3):
        check(hb.describe_plan(shapes, prec, cus=cus), expected(shapes, prec, cus), len(shapes))
        tile = int(rng.integers(0, 5))
        split = int(rng.integers(1, 9)) if tile < 4 else 1
        check(hb.describe_plan(shapes, prec, tile=tile, split_k=split, cus=cus),
              expected(shapes, prec, cus, tile=tile, split=split), len(shapes))


MAX_DIM = 0x7fffffff - 4096  # HBS_MAX_DIM


@pytest.mark.parametrize("shape", [(MAX_DIM + 1, 4, 4), (4, MAX_DIM + 4, 4),
                                   (4, 4, 0x7fffffff - 3), (0x7fffffff, 4, 4)])
def test_sizes_beyond_the_int_limit_are_rejected(shape):
    """Sizes whose tile or split arithmetic would overflow 32-bit ints get
    HBS_ERR_ARG instead of a plan with too few tiles."""
    for tile in (None, 0, 4):
        with pytest.raises(hb.HbsError) as ei:
            hb.describe_plan([shape], tile=tile, split_k=1 if tile in (None, 4) else 4)
        assert ei.value.code == -1, (shape, tile)


@pytest.mark.parametrize("split", [1, 2, 4, 64])
def test_longest_accepted_k_splits_into_whole_ranges(split):
    """K at the limit (a multiple of 4): the split ranges are computed without
    overflow, so a thin problem still gets exactly its split count of items."""
    K = MAX_DIM // 4 * 4
    for prec in (hb.PREC_FP32, hb.PREC_BF16X3):
        info = hb.describe_plan([(64, 64, K)], precision=prec, tile=1, split_k=split)
        exp = expected([(64, 64, K)], prec, tile=1, split=split)
        assert info["tiles"] == 1
        assert info["work_items"] == exp["work_items"] == split
        assert info["split_tiles"] == (1 if split > 1 else 0)
