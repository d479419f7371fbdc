"""Device hash join (HashMap, core/dev_structs.cuh:36-132; hash_map_build_pred_other,
core/join.cuh:107-135) against the reference's join KATs (core/test/join_tests.cu) and the
oracle's chained HostHashMap restatement."""
import numpy as np
import pytest

from conftest import rng
from gpu_util import say

pytestmark = pytest.mark.gpu
U64MAX = np.uint64(0xFFFFFFFFFFFFFFFF)


def test_join_count_1to1(golap, ctx):
    # join_tests.cu:71-106: build ids 0..499,999, probe 10,000,000 ids i % 500,000 -> all hit
    say("join count 1to1 500K x 10M")
    build = np.arange(500_000, dtype=np.int32)
    probe = (np.arange(10_000_000, dtype=np.int64) % 500_000).astype(np.int32)
    db, dp = ctx.to_device(build), ctx.to_device(probe)
    m = ctx.hashmap(build.size, db, golap.I32)
    try:
        ctx.hashmap_build(m, build.size)
        assert ctx.hashmap_filled(m) == 500_000
        assert ctx.hash_join_count(m, dp, golap.I32, probe.size) == 10_000_000
    finally:
        ctx.free_hashmap(m)
        ctx.free(db)
        ctx.free(dp)


def test_join_materialize_rids(golap, ctx):
    # join_tests.cu:109-145: 5,000 build rows, 10,000 probes -> 10,000 matches, rid per probe
    say("join rids 5K x 10K")
    r = rng(2)
    build = r.permutation(np.arange(1, 5001, dtype=np.int32) * 7)
    probe = build[r.integers(0, 5000, 10_000)]
    db, dp = ctx.to_device(build), ctx.to_device(probe)
    out = ctx.alloc(probe.size * 8)
    m = ctx.hashmap(build.size, db, golap.I32)
    try:
        ctx.hashmap_build(m, build.size)
        ctx.hashmap_probe(m, dp, golap.I32, probe.size, out)
        rid = ctx.to_host(out, probe.size, np.uint64)
        assert np.array_equal(build[rid.astype(np.int64)], probe)
    finally:
        ctx.free_hashmap(m)
        for p in (db, dp, out):
            ctx.free(p)


@pytest.mark.parametrize("n_dim", [2556, 200_000])
def test_build_with_predicates_matches_oracle(orc, golap, ctx, n_dim):
    r = rng(n_dim)
    keys = r.permutation(n_dim).astype(np.int32) * 3 + 19920101
    region = r.integers(0, 5, n_dim).astype(np.uint8)
    year = r.integers(1992, 1999, n_dim).astype(np.int16)
    probe = r.integers(19920101, 19920101 + 3 * n_dim + 50, 100_000).astype(np.int32)
    say("build with predicates", n_dim)
    dk, dr, dy, dp = (ctx.to_device(a) for a in (keys, region, year, probe))
    out = ctx.alloc(probe.size * 8)
    m = ctx.hashmap(n_dim, dk, golap.I32)
    try:
        ctx.hashmap_build(m, n_dim, [(dr, golap.U8, 1, 1), (dy, golap.I16, 1995, 1997)])
        ok = (region == 1) & (year >= 1995) & (year <= 1997)
        assert ctx.hashmap_filled(m) == int(ok.sum())
        ctx.hashmap_probe(m, dp, golap.I32, probe.size, out)
        rid = ctx.to_host(out, probe.size, np.uint64)
        # oracle: chained map over the qualifying rows; probe -> rid or UINT64_MAX
        pos = {int(k): i for i, k in enumerate(keys) if ok[i]}
        want = np.array([pos.get(int(k), U64MAX) for k in probe], dtype=np.uint64)
        assert np.array_equal(rid, want)
        s, cnt = orc.q1_plan(keys, [(region, 1, 1), (year, 1995, 1997)], probe,
                             np.ones(probe.size, np.uint8), np.ones(probe.size, np.uint8),
                             np.ones(probe.size, np.int32), (0, 255), (0, 255))
        assert cnt == int((want != U64MAX).sum())
    finally:
        ctx.free_hashmap(m)
        for p in (dk, dr, dy, dp, out):
            ctx.free(p)


def test_capacity_exceeded_reports_error(golap, ctx):
    keys = np.arange(1, 101, dtype=np.int32)
    dk = ctx.to_device(keys)
    m = ctx.hashmap(10, dk, golap.I32)  # room for 10 rows, 100 qualify
    try:
        with pytest.raises(golap.GolapError) as e:
            ctx.hashmap_build(m, keys.size)
        assert e.value.status == golap.ECAPACITY
    finally:
        ctx.free_hashmap(m)
        ctx.free(dk)


def test_reserved_key_rejected(golap, ctx):
    keys = np.array([5, -2**31, 7], dtype=np.int32)
    dk = ctx.to_device(keys)
    m = ctx.hashmap(3, dk, golap.I32)
    try:
        with pytest.raises(golap.GolapError) as e:
            ctx.hashmap_build(m, keys.size)
        assert e.value.status == golap.EINVAL
    finally:
        ctx.free_hashmap(m)
        ctx.free(dk)
