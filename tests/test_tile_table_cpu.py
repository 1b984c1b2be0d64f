"""The tile table and the tuner's bookkeeping, without a GPU: the committed
table is valid for the kernels' candidate lists (the FP64 candidate after
the geometries is the per-wave tile kernel), out-of-range rules are
refused, and the tuner started from a table reproduces it."""
import os
import sys

import pytest

import block_amd

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def test_committed_table_is_valid():
    rules = block_amd.read_tile_table()
    assert rules
    nfp, nsp = block_amd.tile_candidates(block_amd.FP64), block_amd.tile_candidates(block_amd.SPLIT_BF16)
    # FP64 lists one geometry more than split precision (the 384-wide
    # tiles) and, after its geometries, the per-wave tile kernel
    assert nfp == nsp + 2
    for m_lo, m_hi, n_lo, n_hi, fp, sp in rules:
        assert 1 <= m_lo <= m_hi and 1 <= n_lo <= n_hi
        assert -1 <= fp < nfp and -1 <= sp < nsp
    block_amd.set_tile_table(rules)  # accepted as it is


def test_out_of_range_rules_are_refused():
    nfp, nsp = block_amd.tile_candidates(block_amd.FP64), block_amd.tile_candidates(block_amd.SPLIT_BF16)
    try:
        block_amd.set_tile_table([(33, 64, 33, 64, nfp - 1, -1)])  # the per-wave kernel
        for bad in [(33, 64, 33, 64, nfp, -1), (33, 64, 33, 64, -1, nsp), (64, 33, 1, 8, -1, -1),
                    (33, 64, 33, 64, -2, -1)]:
            with pytest.raises(block_amd.BlockError):
                block_amd.set_tile_table([bad])
    finally:
        block_amd.set_tile_table(block_amd.read_tile_table())


def test_tuner_start_round_trip():
    E = bench.TUNE_EDGES
    buckets = [(1, 1), (1, 0), (2, 2)]
    start = [(E[1][0], E[1][1], E[1][0], E[1][1], 9, 3),    # a tuned bucket
             (E[2][0], E[2][1], E[2][0], E[2][1], -1, 1),   # a tuned bucket, split only
             (E[3][0], E[3][1], E[2][0], E[2][1], 1, 1)]    # not a tuned bucket: kept
    choice, keep = bench.tuner_start(buckets, start)
    assert choice == {(1, 1): [9, 3], (1, 0): [-1, -1], (2, 2): [-1, 1]}
    assert keep == [start[2]]
    assert sorted(bench.tuner_rules(choice, keep)) == sorted(start)
    # a bucket back on the model in both precisions gives no rule
    choice[(1, 1)] = [-1, -1]
    assert sorted(bench.tuner_rules(choice, keep)) == sorted(start[1:])
