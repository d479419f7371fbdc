"""Pins the oracle's scan + filter + group-by (oracle_scan_groupby, the checker of every k_scan and
serving test) against a numpy brute force on small tables.

The oracle's contract: a row qualifies when every predicate col in [lo, hi] holds (inclusive, as
the range predicates of usecases/ssb/queries/01.cuh:26-54); its group is the mixed-radix index of
k_i = floor((col_i - base_i) / div_i), first key most significant, each k_i in [0, card_i) (the
grouping of group_by_agg_2col/4col, core/group.cuh:14-36, over computed attributes); SUM(value)
wraps as int64, COUNT counts rows; groups come back in index order with key tuples out_lo_i + k_i.
The brute force below computes the same thing a different way: per-row key tuples, np.unique
over the tuples (lexicographic order = index order), np.add.at for the wrapping sums."""
import numpy as np
import pytest


def brute(cols, preds, keys, val_col):
    n = cols[0].size
    ok = np.ones(n, bool)
    for c, lo, hi in preds:
        v = cols[c].astype(np.int64)
        ok &= (v >= lo) & (v <= hi)
    rows = np.nonzero(ok)[0]
    ks = []
    for c, base, div, card, out_lo in keys:
        d = cols[c][rows].astype(object) - base  # exact, also for int64 extremes
        if any(x < 0 for x in d):
            return None
        k = np.array([x // div for x in d], dtype=object)
        if any(x >= card for x in k):
            return None
        ks.append(np.array(k, np.int64))
    if not keys:
        ks = [np.zeros(rows.size, np.int64)]
    if rows.size == 0:
        return np.zeros((0, len(keys)), np.int32), np.zeros(0, np.int64), np.zeros(0, np.uint64)
    tup = np.stack(ks, axis=1)
    uniq, inv = np.unique(tup, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    sums = np.zeros(uniq.shape[0], np.int64)
    if val_col >= 0:
        vals = cols[val_col][rows].astype(np.int64)
        with np.errstate(over="ignore"):
            np.add.at(sums, inv, vals)
    counts = np.bincount(inv, minlength=uniq.shape[0]).astype(np.uint64)
    out_lo = np.array([k[4] for k in keys], np.int64) if keys else np.zeros(1, np.int64)
    gk = (uniq + out_lo).astype(np.int32)
    if not keys:
        gk = gk[:, :0]
    return gk, sums, counts


def check(orc, cols, preds, keys, val_col, threads=(1, 3)):
    want = brute(cols, preds, keys, val_col)
    for t in threads:
        got = orc.scan_groupby(cols, preds, keys, val_col, threads=t)
        if want is None:
            assert got is None
            continue
        assert got is not None
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]) and \
            np.array_equal(got[2], want[2]), t
    return want


def test_scan_groupby_trip_like_templates(orc):
    """A small trip-style table: time range, zone x hour and passenger x zone keys, fare sums."""
    rng = np.random.default_rng(71)
    n = 20000
    ts = np.sort(rng.integers(1_700_000_000, 1_700_000_000 + 30 * 86400, n)).astype(np.int64)
    pas = rng.integers(1, 7, n).astype(np.uint8)
    zone = rng.integers(1, 266, n).astype(np.uint16)
    fare = rng.integers(250, 50001, n).astype(np.int32)
    cols = [ts, pas, zone, fare]
    t0 = int(ts[0])
    # zone x hour over a time range: div > 1 (3600), a key at card - 1
    hours = int((ts[-1] - t0) // 3600) + 1
    w = check(orc, cols, [(0, t0 + 5 * 3600, t0 + 200 * 3600)], [(2, 1, 1, 265, 1), (0, t0, 3600, hours, 0)], 3)
    assert w[0].shape[0] > 100
    # passenger x zone, COUNT only
    check(orc, cols, [(1, 2, 5)], [(1, 1, 1, 6, 1), (2, 1, 1, 265, 1)], -1)
    # the last zone as the only key value kept: k = card - 1
    w = check(orc, cols, [(2, 265, 265)], [(2, 1, 1, 265, 1)], 3)
    assert w[0].shape[0] == 1 and w[0][0, 0] == 265


def test_scan_groupby_negative_offsets_and_signed_keys(orc):
    rng = np.random.default_rng(72)
    n = 5000
    a = rng.integers(-100, 101, n).astype(np.int8)
    b = rng.integers(-30000, 30001, n).astype(np.int16)
    v = rng.integers(-2**31, 2**31, n).astype(np.int32)
    cols = [a, b, v]
    # negative base: signed key values shifted into [0, card)
    check(orc, cols, [], [(0, -100, 1, 201, -100)], 2)
    # div > 1 over a negative range, floor division of the offset
    check(orc, cols, [(1, -20000, 20000)], [(1, -20000, 7, 40001 // 7 + 1, 0), (0, -100, 50, 5, 0)], 2)
    # a qualifying row below the base is an error (-1 -> None), a filtered one is not
    check(orc, cols, [(0, -100, 100)], [(0, -50, 1, 151, 0)], 2)  # rows < -50 qualify: error
    check(orc, cols, [(0, -50, 100)], [(0, -50, 1, 151, 0)], 2)   # filtered: fine
    # a key past card - 1 is an error
    check(orc, cols, [], [(0, -100, 1, 200, 0)], 2)


def test_scan_groupby_empty_and_wrapping(orc):
    rng = np.random.default_rng(73)
    n = 3000
    k = rng.integers(0, 10, n).astype(np.int32)
    big = np.where(rng.random(n) < 0.5, np.int64(2**62), np.int64(-2**63 + 5)).astype(np.int64)
    cols = [k, big]
    w = check(orc, cols, [(0, 20, 30)], [(0, 0, 1, 10, 0)], 1)  # nothing qualifies
    assert w[0].shape[0] == 0
    w = check(orc, cols, [], [(0, 0, 1, 10, 0)], 1)  # sums wrap as int64
    assert w[0].shape[0] == 10
    w = check(orc, cols, [(0, 3, 3)], [], 1)  # no key: one group
    assert w[1].size == 1 and int(w[2][0]) == int((k == 3).sum())


def test_scan_groupby_shared_accumulators(orc):
    """A domain large enough that the oracle's threads share one accumulator array (atomics)."""
    rng = np.random.default_rng(74)
    n = 40000
    z = rng.integers(0, 2000, n).astype(np.int32)
    h = rng.integers(0, 3000, n).astype(np.int32)
    v = rng.integers(0, 1 << 20, n).astype(np.int32)
    cols = [z, h, v]
    check(orc, cols, [(2, 1000, 1 << 19)], [(0, 0, 1, 2000, 0), (1, 0, 1, 3000, 0)], 2, threads=(1, 8))
