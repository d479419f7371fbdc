"""Fused Q1 plan (query1 + pipeline_q1_lineorder, usecases/ssb/queries/01.hpp:53-449,
01.cuh:26-54) on HBM-resident compressed chunks, bit-exact against the oracle's CPU path
(decode + 01inmem.hpp plan) and against the reference's logged SSB result."""
import numpy as np
import pytest

from gpu_util import q1_dataset, say

pytestmark = pytest.mark.gpu


def upload_q1(ctx, golap, ds, **kw):
    cols = []
    hosts = []
    for name in ("date", "disc", "qty", "price"):
        hc = golap.HostColumn.encode(ds[name], **kw)
        hosts.append(hc)
        cols.append(ctx.upload(hc))
    return cols, hosts


def date_map(ctx, golap, ds, year):
    dk = ctx.to_device(ds["d_key"])
    dy = ctx.to_device(ds["d_year"])
    m = ctx.hashmap(ds["d_key"].size, dk, golap.I32)
    ctx.hashmap_build(m, ds["d_key"].size, [(dy, golap.I16, year, year)])
    return m, [dk, dy]


def release(ctx, cols, m, bufs):
    for c in cols:
        ctx.free_column(c)
    ctx.free_hashmap(m)
    for b in bufs:
        ctx.free(b)


PLANS = [  # (year, discount range, quantity range, price range)
    (1995, (4, 6), (30, 50), None),       # the workload's plan
    (1993, (1, 3), (0, 24), None),        # SSB Q1.1 constants
    (1994, (5, 7), (26, 35), (1_000_000, 5_000_000)),
    (1990, (0, 10), (1, 50), None),       # no date qualifies
    (1992, (0, 255), (0, 255), None),     # every fact row passes the fact predicates
]


@pytest.mark.parametrize("plan", range(len(PLANS)))
@pytest.mark.parametrize("codec,mm,page_rows", [(1, 16, 4096), (1, 4, 1024), (0, 4, 2048), (255, 4, 4096),
                                                 (1, 4, 4096), (1, 4, 8192), ("typed", 4, 4096)])
def test_q1_matches_oracle(orc, golap, ctx, plan, codec, mm, page_rows):
    n = 6_000_000 if (plan == 0 and codec == 1 and page_rows == 4096) else 1_000_003
    ds = q1_dataset(orc, n)
    year, drng, qrng, prng = PLANS[plan]
    prng = prng or (-2**31, 2**31 - 1)
    say("q1 plan", plan, "codec", codec, "min_match", mm, "page_rows", page_rows, "rows", n)
    unit = 1
    if codec == "typed":  # LZ4 in whole units of the element width (date and price: 4 bytes)
        codec, unit = 1, 0
    cols, hosts = upload_q1(ctx, golap, ds, codec=codec, page_rows=page_rows, chunk_rows=262144,
                            lz4_min_match=mm, lz4_unit=unit)
    m, bufs = date_map(ctx, golap, ds, year)
    try:
        r = ctx.q1_run(*cols, m, drng, qrng, prng)
        want, wcnt = orc.q1_plan(ds["d_key"], [(ds["d_year"], year, year)], ds["date"], ds["disc"],
                                 ds["qty"], ds["price"], drng, qrng, prng, threads=4)
        assert (r.sum, r.matches) == (want, wcnt)
        # the CPU path over the very same compressed pages
        arrs = [h.arrays() for h in hosts]
        d = arrs[0]["desc"]
        cpu, ccnt = orc.q1_compressed(ds["d_key"], [(ds["d_year"], year, year)], arrs, n, d.page_rows,
                                      d.chunk_rows, 0, d.n_chunks, drng, qrng, prng, threads=4)
        assert (cpu, ccnt) == (want, wcnt)
        assert r.chunks_run == d.n_chunks and r.chunks_pruned == 0
        assert r.comp_bytes == sum(a["desc"].payload_bytes for a in arrs)
    finally:
        release(ctx, cols, m, bufs)


def test_q1_chunk_range_and_empty_range(orc, golap, ctx):
    ds = q1_dataset(orc, 1_500_000, seed=3)
    cols, _ = upload_q1(ctx, golap, ds, codec=1, page_rows=4096, chunk_rows=262144, lz4_min_match=16)
    m, bufs = date_map(ctx, golap, ds, 1995)
    try:
        r = ctx.q1_run(*cols, m, (4, 6), (30, 50), chunk_begin=2, chunk_end=5)
        sl = slice(2 * 262144, 5 * 262144)
        want, cnt = orc.q1_plan(ds["d_key"], [(ds["d_year"], 1995, 1995)], ds["date"][sl], ds["disc"][sl],
                                ds["qty"][sl], ds["price"][sl], (4, 6), (30, 50))
        assert (r.sum, r.matches) == (want, cnt)
        e = ctx.q1_run(*cols, m, (4, 6), (30, 50), chunk_begin=3, chunk_end=3)
        assert (e.sum, e.matches, e.chunks_run) == (0, 0, 0)
    finally:
        release(ctx, cols, m, bufs)


def test_q1_zone_map_pruning(orc, golap, ctx):
    """MetaChecker::add_minmax + skip_step (metadata.cuh:645-658, 01.hpp:368-378): on a
    date-sorted table, pruning by a date range gives the same sum as no pruning."""
    ds = q1_dataset(orc, 2_000_000, seed=8)
    order = np.argsort(ds["date"], kind="stable")
    for k in ("date", "disc", "qty", "price"):
        ds[k] = np.ascontiguousarray(ds[k][order])
    cols, hosts = upload_q1(ctx, golap, ds, codec=1, page_rows=4096, chunk_rows=65536, lz4_min_match=16)
    m, bufs = date_map(ctx, golap, ds, 1995)
    nch = ctx.column_desc(cols[0]).n_chunks
    flags = ctx.alloc(nch * 2)
    f2 = ctx.alloc(nch * 2)
    try:
        ctx.minmax_check(cols[0], 19950101, 19951231, flags)
        ctx.minmax_check(cols[1], 4, 6, f2)
        ctx.combine_and(flags, f2, nch)
        got = ctx.to_host(flags, nch, np.uint16)
        a = hosts[0].arrays()
        want = orc.minmax_check(a["chunk_min"], a["chunk_max"], 19950101, 19951231)
        want &= orc.minmax_check(hosts[1].arrays()["chunk_min"], hosts[1].arrays()["chunk_max"], 4, 6)
        assert np.array_equal(got, want)
        assert 0 < want.sum() < nch
        pruned = ctx.q1_run(*cols, m, (4, 6), (30, 50), dev_chunk_mask=flags)
        full = ctx.q1_run(*cols, m, (4, 6), (30, 50))
        assert (pruned.sum, pruned.matches) == (full.sum, full.matches)
        assert pruned.chunks_pruned == nch - int(want.sum())
        cpu, cnt = orc.q1_plan(ds["d_key"], [(ds["d_year"], 1995, 1995)], ds["date"], ds["disc"], ds["qty"],
                               ds["price"], (4, 6), (30, 50))
        assert (full.sum, full.matches) == (cpu, cnt)
    finally:
        release(ctx, cols, m, bufs + [flags, f2])


Q1_SSB = [  # (date predicate columns, discount, quantity); 01.hpp:451-490
    (["year"], [(1993, 1993)], (1, 3), (0, 24)),
    (["yearmonthnum"], [(199401, 199401)], (4, 6), (26, 35)),
    (["weeknuminyear", "year"], [(6, 6), (1994, 1994)], (5, 7), (26, 35)),
]
SF10_GOLDEN = [4212228344835, 900952032730, 246824540411]  # query_sf10_export.csv


@pytest.fixture(scope="module")
def ssb10(orc):
    return orc.ssb_generate(10)


@pytest.mark.parametrize("q", [0, 1, 2])
def test_ssb_sf10_q1_reproduces_reference_log(golap, ctx, ssb10, q):
    lo, d, _, _, _ = ssb10
    say("ssb sf10 q1." + str(q + 1), "rows", lo["orderdate"].size)
    price = lo["extendedprice"]
    assert price.max() < 2**31
    cols = []
    for arr in (lo["orderdate"], lo["discount"], lo["quantity"], price):
        cols.append(ctx.upload(golap.HostColumn.encode(arr, codec=1, page_rows=4096, chunk_rows=262144,
                                                       lz4_min_match=4)))
    bufs = []
    dk = ctx.to_device(d["key"])
    bufs.append(dk)
    m = ctx.hashmap(d["key"].size, dk, golap.I32)
    preds = []
    names, rngs, drng, qrng = Q1_SSB[q]
    for nm, (lo_, hi_) in zip(names, rngs):
        arr = d[nm]
        p = ctx.to_device(arr)
        bufs.append(p)
        preds.append((p, golap.I16 if arr.dtype == np.int16 else golap.I32, lo_, hi_))
    try:
        ctx.hashmap_build(m, d["key"].size, preds)
        r = ctx.q1_run(*cols, m, drng, qrng)
        assert r.sum == SF10_GOLDEN[q]
    finally:
        release(ctx, cols, m, bufs)


@pytest.mark.parametrize("years,page_rows", [((1995, 1995), 4096), ((1992, 1998), 4096), ((1993, 1994), 2048)])
def test_q1_hash_probe_modes_match_oracle(orc, golap, ctx, years, page_rows):
    """A date dimension with keys spread too wide for the membership bitmap: the date probe is
    the open-addressing table, pinned in LDS (one year qualifies: 365 keys) or in HBM (all seven
    years: 2,556 keys), on min-match-4 LZ4I pages (the one-wave-per-tile kernel, which probes the
    date after the quantity and discount tests)."""
    n = 1_500_017
    r = np.random.default_rng(77)
    key = (np.arange(2556, dtype=np.int64) * 997 + 13).astype(np.int32)  # range 2.5M > 2^18
    year = (1992 + np.arange(2556) // 366).astype(np.int16)
    ds = dict(d_key=key, d_year=year, date=key[r.integers(0, 2556, n)],
              disc=r.integers(0, 11, n).astype(np.uint8), qty=r.integers(1, 51, n).astype(np.uint8),
              price=r.integers(90_000, 10_500_001, n).astype(np.int32))
    say("q1 hash probe", years, "page_rows", page_rows)
    cols, hosts = upload_q1(ctx, golap, ds, codec=1, page_rows=page_rows, chunk_rows=262144, lz4_min_match=4)
    dk, dy = ctx.to_device(key), ctx.to_device(year)
    m = ctx.hashmap(key.size, dk, golap.I32)
    ctx.hashmap_build(m, key.size, [(dy, golap.I16, years[0], years[1])])
    try:
        res = ctx.q1_run(*cols, m, (4, 6), (30, 50))
        want, cnt = orc.q1_plan(key, [(year, years[0], years[1])], ds["date"], ds["disc"], ds["qty"], ds["price"],
                                (4, 6), (30, 50), (-2**31, 2**31 - 1), threads=4)
        assert (res.sum, res.matches) == (want, cnt)
    finally:
        release(ctx, cols, m, [dk, dy])
