"""Chunk metadata beyond min/max: EqHistogram and BloomMeta (core/metadata.cuh:225-537) built on
the device per chunk (prepare_compressed's fill, core/apps.cuh:212-224) and checked per chunk
(check_hist / check_bloom). Histograms, filters and flags must equal the oracle's restatement
bit for bit; the checks must never prune a chunk that holds a qualifying value."""
import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu

TYPES = [("I8", np.int8), ("U16", np.uint16), ("I32", np.int32), ("I64", np.int64)]


def column(rng, dtype, n, chunk):
    info = np.iinfo(dtype)
    lo, hi = max(info.min, -50_000), min(info.max, 50_000)
    v = rng.integers(lo, hi, n).astype(dtype)
    v[chunk * 2:chunk * 3] = dtype(7)                 # a constant chunk
    v[chunk * 4:chunk * 5] = np.sort(v[chunk * 4:chunk * 5])
    return v


@pytest.mark.parametrize("tname,dtype", TYPES)
@pytest.mark.parametrize("buckets", [1, 7, 64, 4096])
def test_eqhist_equals_oracle(orc, golap, ctx, tname, dtype, buckets):
    rng = np.random.default_rng(buckets)
    chunk = 3000
    v = column(rng, dtype, chunk * 9 + 123, chunk)  # ragged last chunk
    dv = ctx.to_device(v)
    h = ctx.eqhist_build(dv, getattr(golap, tname), v.size, chunk, buckets)
    nch = -(-v.size // chunk)
    flags = ctx.alloc(nch * 2)
    try:
        d = ctx.chunkmeta_desc(h)
        mins, maxs, hist = orc.eqhist_build(v.astype(np.int64), chunk, buckets)
        assert d.n_chunks == nch and d.buckets == buckets
        assert np.array_equal(ctx.to_host(d.mins, nch, np.int64), mins)
        assert np.array_equal(ctx.to_host(d.maxs, nch, np.int64), maxs)
        assert np.array_equal(ctx.to_host(d.hist, nch * buckets, np.uint64).reshape(nch, buckets), hist)
        ch = [v[c * chunk:(c + 1) * chunk].astype(np.int64) for c in range(nch)]
        for _ in range(40):
            lo = int(rng.integers(int(v.min()) - 10, int(v.max()) + 10))
            hi = lo + int(rng.integers(0, 3000))
            ctx.eqhist_check(h, lo, hi, flags)
            got = ctx.to_host(flags, nch, np.uint16)
            assert np.array_equal(got, orc.eqhist_check(mins, maxs, hist, lo, hi)), (lo, hi)
            has = np.array([((x >= lo) & (x <= hi)).any() for x in ch])
            assert not np.any(has & (got == 0))
        for lo, hi in ((7, 7), (int(maxs[0]), int(maxs[0])), (int(mins[3]), int(mins[3]))):  # edges
            ctx.eqhist_check(h, lo, hi, flags)
            assert np.array_equal(ctx.to_host(flags, nch, np.uint16), orc.eqhist_check(mins, maxs, hist, lo, hi))
    finally:
        ctx.free(flags)
        ctx.chunkmeta_free(h)
        ctx.free(dv)


@pytest.mark.parametrize("tname,dtype", TYPES)
@pytest.mark.parametrize("p,m", [(0.01, 1024), (0.1, 64), (0.001, 4096)])
def test_bloommeta_equals_oracle(orc, golap, ctx, tname, dtype, p, m):
    rng = np.random.default_rng(int(m))
    chunk = 700
    v = column(rng, dtype, chunk * 11 + 5, chunk)
    dv = ctx.to_device(v)
    h = ctx.bloommeta_build(dv, getattr(golap, tname), v.size, chunk, p, m)
    nch = -(-v.size // chunk)
    flags = ctx.alloc(nch * 2)
    try:
        d = ctx.chunkmeta_desc(h)
        n, k = orc.bloom_params(p, m)
        assert (d.bloom_n, d.bloom_k, d.bloom_bytes, d.n_chunks) == (n, k, m, nch)
        state = orc.bloom_build(v.astype(np.int64), chunk, m, k)
        assert np.array_equal(ctx.to_host(d.bloom_state, nch * m // 4, np.uint32), state)
        values = list(rng.choice(v, 20).astype(np.int64)) + list(rng.integers(-60_000, 60_000, 20))
        for x in values:
            ctx.bloommeta_check(h, int(x), flags)
            got = ctx.to_host(flags, nch, np.uint16)
            xt = int(np.array([x]).astype(dtype).astype(np.int64)[0])  # the value as the column type holds it
            assert np.array_equal(got, orc.bloom_check(state, nch, m, k, xt)), x
            has = np.array([(v[c * chunk:(c + 1) * chunk] == dtype(xt)).any() for c in range(nch)])
            assert not np.any(has & (got == 0))
        say(tname, p, m, "k", k)
    finally:
        ctx.free(flags)
        ctx.chunkmeta_free(h)
        ctx.free(dv)


def test_chunkmeta_rejects_bad_parameters(golap, ctx):
    dv = ctx.to_device(np.arange(100, dtype=np.int32))
    try:
        for b in (0, 8193):
            with pytest.raises(golap.GolapError):
                ctx.eqhist_build(dv, golap.I32, 100, 10, b)
        for p, m in ((0.0, 64), (1.0, 64), (0.01, 66), (0.01, 1 << 17)):
            with pytest.raises(golap.GolapError):
                ctx.bloommeta_build(dv, golap.I32, 100, 10, p, m)
    finally:
        ctx.free(dv)
