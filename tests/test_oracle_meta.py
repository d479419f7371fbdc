"""The oracle's EqHistogram / BloomMeta restatement (core/metadata.cuh:225-537) pinned by
hand-computed buckets, the reference's parameter formulas worked by hand, and the properties
the pruning relies on (no chunk holding a qualifying value is pruned)."""
import numpy as np

import oracle as orc


def test_eqhist_known_buckets():
    # one chunk of 0..99, 10 buckets of width 9.9: bucket b holds [9.9 b, 9.9 (b + 1)), i.e.
    # 0..9, 10..19, ..., 80..89, and 90..98 plus the maximum 99 (10.0 clamped to 9)
    mins, maxs, hist = orc.eqhist_build(np.arange(100), 100, 10)
    assert (mins[0], maxs[0]) == (0, 99)
    assert hist[0].tolist() == [10] * 10
    # 0..20, 4 buckets of width 5: 0..4, 5..9, 10..14, 15..19 plus 20
    _, _, h = orc.eqhist_build(np.arange(21), 21, 4)
    assert h[0].tolist() == [5, 5, 5, 6]


def test_eqhist_constant_chunk_and_check_rules():
    v = np.array([7] * 50 + list(range(10, 60)), np.int64)  # chunk 0 constant 7, chunk 1 10..59
    mins, maxs, hist = orc.eqhist_build(v, 50, 8)
    assert hist[0][0] == 50 and hist[0][1:].sum() == 0
    chk = lambda lo, hi: orc.eqhist_check(mins, maxs, hist, lo, hi).tolist()
    # chunk 1: min 10, width 49/8. A range ending less than one bucket width below the minimum
    # maps to bucket -0.x, which the reference's int64 conversion truncates to 0: kept
    assert chk(7, 7) == [1, 1]
    assert chk(0, 3) == [0, 0]   # (3 - 10) * 8 / 49 = -1.14 -> -1: below every bucket
    assert chk(59, 100) == [0, 1]  # lo == max: the last bucket
    assert chk(60, 100) == [0, 0]
    assert chk(-5, 9) == [1, 1]


def test_eqhist_never_prunes_a_chunk_with_a_qualifying_value():
    rng = np.random.default_rng(3)
    v = rng.integers(-1000, 1000, 64 * 1000)
    v[64 * 5:64 * 6] = 42  # a constant chunk
    mins, maxs, hist = orc.eqhist_build(v, 64, 16)
    ch = v.reshape(-1, 64)
    for _ in range(200):
        lo = int(rng.integers(-1100, 1100))
        hi = lo + int(rng.integers(0, 50))
        got = orc.eqhist_check(mins, maxs, hist, lo, hi)
        has = ((ch >= lo) & (ch <= hi)).any(axis=1)
        assert not np.any(has & (got == 0)), (lo, hi)
        # a chunk whose range misses [lo, hi] by at least a bucket width is pruned
        width = (maxs - mins) / 16.0
        assert not np.any((got == 1) & (mins - hi >= width) & (maxs > mins))
        assert not np.any((got == 1) & (lo > maxs))


def test_bloom_params_by_hand():
    # m = 1024 bytes: mbits 8192, n = ceil(8192 * ln(0.5) * ln(2) / ln(0.01)) = 855, k = ceil(ln2 * 8192 / 855) = 7
    assert orc.bloom_params(0.01, 1024) == (855, 7)
    # m = 64 bytes, p = 0.1: n = ceil(512 * 0.48045 / 2.302585) = 107, k = ceil(0.6931 * 512 / 107) = 4
    assert orc.bloom_params(0.1, 64) == (107, 4)


def test_bloom_no_false_negatives_and_rate():
    n, k = orc.bloom_params(0.01, 1024)
    rng = np.random.default_rng(4)
    rows = 855
    v = rng.integers(-2**40, 2**40, rows * 20)
    state = orc.bloom_build(v, rows, 1024, k)
    assert state.size == 20 * 1024 // 4
    for c in (0, 7, 19):
        for x in v[c * rows:(c + 1) * rows][:50]:
            assert orc.bloom_check(state, 20, 1024, k, int(x))[c] == 1
    # values never inserted: the false positive rate stays near p at the design load
    fp = sum(int(orc.bloom_check(state, 20, 1024, k, int(x)).sum()) for x in rng.integers(2**41, 2**42, 500))
    assert fp / (500 * 20) < 0.03
    # each inserted value sets at most k bits: a one-value chunk has popcount <= k
    s1 = orc.bloom_build(np.array([123], np.int64), 1, 1024, k)
    assert 1 <= int(np.unpackbits(s1.view(np.uint8)).sum()) <= k
