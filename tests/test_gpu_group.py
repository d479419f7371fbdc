"""Group-by (group_by_agg_2col/4col, core/group.cuh:14-36; HashAggregate::add,
core/dev_structs.cuh:209-244) against the reference KATs of core/test/group_by_tests.cu and the
oracle, for every aggregation mode (LDS partial tables with spill, global atomics)."""
import numpy as np
import pytest

from conftest import rng
from gpu_util import say

pytestmark = pytest.mark.gpu
MODES = [0, 1, 2]


@pytest.mark.parametrize("mode", MODES)
def test_group_by_agg_columns_kat(golap, ctx, mode):
    # group_by_tests.cu:75-132: rows i, group i % 64, each group's sum == rows / 64 (value 1)
    n = 1 << 26
    say("group by 2^26 rows x 64 groups, mode", mode)
    keys = (np.arange(n, dtype=np.int64) % 64).astype(np.int32)
    dk = ctx.to_device(keys)
    dv = ctx.to_device(np.ones(n, np.int32))
    try:
        k, s, c = ctx.group_by([(dk, golap.I32, 0, 64)], dv, golap.I32, n, mode)
        assert np.array_equal(k[:, 0], np.arange(64))
        assert np.all(s == n // 64) and np.all(c == n // 64)
    finally:
        ctx.free(dk)
        ctx.free(dv)


@pytest.mark.parametrize("mode", MODES)
def test_large_group_four_key_columns(golap, ctx, mode):
    # group_by_tests.cu:135-196: 4M rows, 4 key columns, 32 groups
    n = 4 << 20
    i = np.arange(n, dtype=np.int64)
    cols = [(i % 2).astype(np.int32), ((i // 2) % 2).astype(np.int32), ((i // 4) % 2).astype(np.int32),
            ((i // 8) % 4).astype(np.int32)]
    val = (i % 1000).astype(np.int64) - 500
    dcols = [ctx.to_device(c) for c in cols]
    dv = ctx.to_device(val)
    try:
        k, s, c = ctx.group_by([(d, golap.I32, 0, card) for d, card in zip(dcols, (2, 2, 2, 4))], dv, golap.I64,
                               n, mode)
        assert k.shape == (32, 4)
        gid = ((cols[0] * 2 + cols[1]) * 2 + cols[2]) * 4 + cols[3]
        want_s = np.bincount(gid, weights=val, minlength=32).astype(np.int64)
        assert np.array_equal(s, want_s) and np.all(c == n // 32)
    finally:
        for d in dcols + [dv]:
            ctx.free(d)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("card", [(7, 1000), (265, 8784)])
def test_group_by_matches_oracle(orc, golap, ctx, mode, card):
    r = rng(card[1])
    n = 3_000_000
    a = r.integers(0, card[0], n).astype(np.int32) + 1992
    b = r.integers(0, card[1], n).astype(np.int32)
    v = r.integers(-2**31, 2**31 - 1, n).astype(np.int64)
    say("group by vs oracle", card, "mode", mode)
    da, db, dv = ctx.to_device(a), ctx.to_device(b), ctx.to_device(v)
    try:
        k, s, c = ctx.group_by([(da, golap.I32, 1992, card[0]), (db, golap.I32, 0, card[1])], dv, golap.I64, n, mode,
                               max_groups=card[0] * card[1])
        wk, ws, wc = orc.groupby([a, b], v, max_groups=card[0] * card[1])
        assert np.array_equal(k, wk) and np.array_equal(s, ws) and np.array_equal(c, wc)
    finally:
        for d in (da, db, dv):
            ctx.free(d)


def test_group_key_out_of_range(golap, ctx):
    a = np.array([0, 1, 5, 2], np.int32)
    da, dv = ctx.to_device(a), ctx.to_device(np.ones(4, np.int32))
    try:
        with pytest.raises(golap.GolapError) as e:
            ctx.group_by([(da, golap.I32, 0, 4)], dv, golap.I32, 4)
        assert e.value.status == golap.EINVAL
    finally:
        ctx.free(da)
        ctx.free(dv)
