"""Fused star join + group-by (query2/query3 with pipeline_q2/pipeline_q3,
usecases/ssb/queries/02.hpp, 02.cuh:102-193, 03.cuh:118-170) against the oracle's restatement of
the CPU plan (02inmem.hpp:58-206) and the reference's logged SSB Q2.x results."""
import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu


def star_data(golap, n, seed=17, n_cust=20000, n_supp=2000, n_part=14000):
    key, year, _ = golap.gen_date_dim(2556)
    fact = golap.gen_star_fact(seed, n, n, key, n_cust, n_supp, n_part)
    nation = golap.gen_uniform_i32(seed, 20, n_cust, 0, 24)
    region = golap.gen_uniform_i32(seed, 21, n_supp, 0, 4)
    brand = golap.gen_uniform_i32(seed, 22, n_part, 1, 1000)
    return dict(d_key=key, d_year=year.astype(np.int32), c_nation=nation, s_region=region, p_brand=brand,
                c_key=np.arange(n_cust, dtype=np.int32), s_key=np.arange(n_supp, dtype=np.int32),
                p_key=np.arange(n_part, dtype=np.int32), date=fact[0], cust=fact[1], supp=fact[2], part=fact[3],
                rev=fact[4])


def build_map(ctx, golap, key, n, preds, payload):
    dk = ctx.to_device(key)
    bufs = [dk]
    dp = []
    for col, lo, hi in preds:
        d = ctx.to_device(col)
        bufs.append(d)
        dp.append((d, golap.TYPE_OF[col.dtype], lo, hi))
    dpay = None
    if payload is not None:
        dpay = ctx.to_device(np.ascontiguousarray(payload, np.int32))
        bufs.append(dpay)
    m = ctx.hashmap(n, dk, golap.I32)
    ctx.hashmap_build(m, n, dp, dpay)
    return m, bufs


QUERIES = [
    # (supplier region, customer nation, brand range) -> group (year, brand)
    (1, 7, (42, 42)),
    (3, 11, (100, 199)),
    (0, 0, (1, 1000)),
]


@pytest.mark.parametrize("q", range(len(QUERIES)))
@pytest.mark.parametrize("codec,mm,date_codec", [(1, 16, 255), (1, 4, 1), (0, 4, 0), (2, 4, 2)])
@pytest.mark.parametrize("mode", [0, 2])
def test_star_matches_oracle(orc, golap, ctx, q, codec, mm, date_codec, mode):
    n = 2_000_000
    D = star_data(golap, n)
    reg, nat, (blo, bhi) = QUERIES[q]
    say("star", q, "codec", codec, mm, "date codec", date_codec, "mode", mode)
    enc = dict(page_rows=4096, chunk_rows=65536, lz4_min_match=mm)
    cols = {}
    for name in ("date", "cust", "supp", "part", "rev"):
        cols[name] = ctx.upload(golap.HostColumn.encode(D[name], codec=date_codec if name == "date" else codec, **enc))
    maps, bufs = [], []
    for key, n_dim, preds, pay in (
            (D["d_key"], 2556, [], D["d_year"]),
            (D["c_key"], D["c_key"].size, [(D["c_nation"], nat, nat)], None),
            (D["s_key"], D["s_key"].size, [(D["s_region"], reg, reg)], None),
            (D["p_key"], D["p_key"].size, [(D["p_brand"], blo, bhi)], D["p_brand"])):
        m, b = build_map(ctx, golap, key, n_dim, preds, pay)
        maps.append(m)
        bufs += b
    try:
        dims = [(cols["date"], maps[0], 1992, 7), (cols["cust"], maps[1], 0, 0), (cols["supp"], maps[2], 0, 0),
                (cols["part"], maps[3], 1, 1000)]
        r, k, s, c = ctx.star_run(dims, cols["rev"], agg_mode=mode)
        wk, ws, wc = orc.star_groupby(
            [(D["d_key"], [], D["d_year"]), (D["c_key"], [(D["c_nation"], nat, nat)], None),
             (D["s_key"], [(D["s_region"], reg, reg)], None), (D["p_key"], [(D["p_brand"], blo, bhi)], D["p_brand"])],
            [D["date"], D["cust"], D["supp"], D["part"]], D["rev"], threads=4)
        assert np.array_equal(k, wk) and np.array_equal(s, ws) and np.array_equal(c, wc)
        assert r.rows_matched == int(wc.sum()) and r.total_sum == int(ws.sum())
        sel = [1.0, float(np.mean(D["c_nation"] == nat)), float(np.mean(D["s_region"] == reg)),
               float(np.mean((D["p_brand"] >= blo) & (D["p_brand"] <= bhi)))]
        assert r.driver_dim == int(np.argmin(sel))  # the most selective dimension drives
    finally:
        for c_ in cols.values():
            ctx.free_column(c_)
        for m in maps:
            ctx.free_hashmap(m)
        for b in bufs:
            ctx.free(b)


@pytest.mark.parametrize("mm", [4, 16])
def test_star_config3_dimension_sizes_match_oracle(orc, golap, ctx, mm):
    """The star plan at its real dimension sizes: customer 2,000,000, supplier 200,000, part
    1,400,000 rows (key ranges above the exact driver bitmap's 131,072, so the driver filter is the
    blocked Bloom filter) and predicates at 1/5 (supplier region), 1/25 (customer nation) and
    1/1000 (part brand), over 24M fact rows coded as the star workload codes them (date_key
    bit-packed delta, keys and revenue LZ4). Groups, sums and counts equal the oracle's CPU plan."""
    n = 24_000_000
    D = star_data(golap, n, seed=23, n_cust=2_000_000, n_supp=200_000, n_part=1_400_000)
    say("star config-3 dims, rows", n, "min match", mm)
    enc = dict(page_rows=4096, chunk_rows=262144, lz4_min_match=mm)
    cols = {nm: ctx.upload(golap.HostColumn.encode(D[nm], codec=2 if nm == "date" else 1, **enc))
            for nm in ("date", "cust", "supp", "part", "rev")}
    maps, bufs = [], []
    specs = ((D["d_key"], 2556, [], D["d_year"]),
             (D["c_key"], D["c_key"].size, [(D["c_nation"], 7, 7)], None),
             (D["s_key"], D["s_key"].size, [(D["s_region"], 1, 1)], None),
             (D["p_key"], D["p_key"].size, [(D["p_brand"], 42, 42)], D["p_brand"]))
    for key, n_dim, preds, pay in specs:
        m, b = build_map(ctx, golap, key, n_dim, preds, pay)
        maps.append(m)
        bufs += b
    try:
        dims = [(cols["date"], maps[0], 1992, 7), (cols["cust"], maps[1], 0, 0), (cols["supp"], maps[2], 0, 0),
                (cols["part"], maps[3], 1, 1000)]
        wk, ws, wc = orc.star_groupby([(k, p, pay) for k, _, p, pay in specs],
                                      [D["date"], D["cust"], D["supp"], D["part"]], D["rev"], threads=8)
        assert wk.shape[0] > 0
        for mode in (0, 1, 2):
            r, k, s, c = ctx.star_run(dims, cols["rev"], agg_mode=mode)
            say("mode", mode, "groups", r.n_groups, "rows", r.rows_matched, "driver", r.driver_dim,
                "kernel ms", round(r.kernel_ms, 3))
            assert np.array_equal(k, wk) and np.array_equal(s, ws) and np.array_equal(c, wc), mode
            assert r.driver_dim == 3  # part brand (1/1000) drives through the Bloom filter
    finally:
        for c_ in cols.values():
            ctx.free_column(c_)
        for m in maps:
            ctx.free_hashmap(m)
        for b in bufs:
            ctx.free(b)


def test_star_with_zone_map_pruning(orc, golap, ctx):
    """Date-sorted fact: pruning chunks by a date range gives the result of the date predicate."""
    n = 3_000_000
    D = star_data(golap, n, seed=3)
    enc = dict(page_rows=4096, chunk_rows=65536, lz4_min_match=16)
    cols = {nm: ctx.upload(golap.HostColumn.encode(D[nm], codec=255 if nm == "date" else 1, **enc))
            for nm in ("date", "cust", "supp", "part", "rev")}
    # the date map keeps 1995..1996 only; pruning by the same range must not change the result
    maps, bufs = [], []
    for key, n_dim, preds, pay in ((D["d_key"], 2556, [(D["d_year"], 1995, 1996)], D["d_year"]),
                                   (D["s_key"], D["s_key"].size, [(D["s_region"], 2, 2)], None)):
        m, b = build_map(ctx, golap, key, n_dim, preds, pay)
        maps.append(m)
        bufs += b
    nch = ctx.column_desc(cols["date"]).n_chunks
    flags = ctx.alloc(nch * 2)
    try:
        ctx.minmax_check(cols["date"], 19950101, 19961231, flags)
        dims = [(cols["date"], maps[0], 1992, 7), (cols["supp"], maps[1], 0, 0)]
        r1, k1, s1, c1 = ctx.star_run(dims, cols["rev"], dev_chunk_mask=flags)
        r2, k2, s2, c2 = ctx.star_run(dims, cols["rev"])
        assert np.array_equal(k1, k2) and np.array_equal(s1, s2) and np.array_equal(c1, c2)
        assert r1.chunks_pruned > nch // 2
        wk, ws, wc = orc.star_groupby([(D["d_key"], [(D["d_year"], 1995, 1996)], D["d_year"]),
                                       (D["s_key"], [(D["s_region"], 2, 2)], None)],
                                      [D["date"], D["supp"]], D["rev"], threads=4)
        assert np.array_equal(k1, wk) and np.array_equal(s1, ws) and np.array_equal(c1, wc)
    finally:
        for c_ in cols.values():
            ctx.free_column(c_)
        for m in maps:
            ctx.free_hashmap(m)
        for b in bufs + [flags]:
            ctx.free(b)


def test_star_masked_page_ranges_match_oracle(orc, golap, ctx):
    """A mask keeping 45 of 62 chunks of 64 pages (1,024-row pages, 65,536-row chunks): the
    surviving chunks are cut into page ranges sized by their count (23 ranges of 3 pages on 256
    CUs: the last range of a chunk is clamped, ranges past its end are empty); the groups equal the
    oracle's over the surviving chunks' rows."""
    n = 62 * 65536 - 1000
    D = star_data(golap, n, seed=31)
    enc = dict(page_rows=1024, chunk_rows=65536, lz4_min_match=4)
    cols = {nm: ctx.upload(golap.HostColumn.encode(D[nm], codec=1, **enc)) for nm in ("date", "supp", "rev")}
    maps, bufs = [], []
    for key, n_dim, preds, pay in ((D["d_key"], 2556, [], D["d_year"]),
                                   (D["s_key"], D["s_key"].size, [(D["s_region"], 2, 2)], None)):
        m, b = build_map(ctx, golap, key, n_dim, preds, pay)
        maps.append(m)
        bufs += b
    nch = ctx.column_desc(cols["date"]).n_chunks
    keep = np.ones(nch, np.uint16)
    keep[np.random.default_rng(32).choice(nch, nch - 45, replace=False)] = 0
    flags = ctx.to_device(keep)
    say("star masked page ranges: chunks", nch, "kept", int(keep.sum()))
    try:
        dims = [(cols["date"], maps[0], 1992, 7), (cols["supp"], maps[1], 0, 0)]
        r, k, s_, c = ctx.star_run(dims, cols["rev"], dev_chunk_mask=flags)
        assert r.chunks_pruned == nch - 45 and r.chunks_run == 45
        rows = np.repeat(keep.astype(bool), 65536)[:n]
        wk, ws, wc = orc.star_groupby([(D["d_key"], [], D["d_year"]), (D["s_key"], [(D["s_region"], 2, 2)], None)],
                                      [D["date"][rows], D["supp"][rows]], D["rev"][rows], threads=4)
        assert np.array_equal(k, wk) and np.array_equal(s_, ws) and np.array_equal(c, wc)
    finally:
        for c_ in cols.values():
            ctx.free_column(c_)
        for m in maps:
            ctx.free_hashmap(m)
        for b in bufs + [flags]:
            ctx.free(b)


SF10_Q2 = [(1737351738436, 280), (347570089279, 56), (43762644335, 7)]  # query_sf10_export.csv


@pytest.fixture(scope="module")
def ssb10(orc):
    return orc.ssb_generate(10)


@pytest.mark.parametrize("q", [0, 1, 2])
def test_ssb_sf10_q2_reproduces_reference_log(orc, golap, ctx, ssb10, q):
    """Q2.1: p_category = MFGR#12, s_region = AMERICA; Q2.2: p_brand1 in MFGR#2221..MFGR#2228,
    s_region = ASIA; Q2.3: p_brand1 = MFGR#2239, s_region = EUROPE; group by (d_year, p_brand1)."""
    lo, d, p, s, _ = ssb10
    say("ssb sf10 q2." + str(q + 1), "rows", lo["orderdate"].size)
    n_parts, n_supp = p["category"].size, s["region"].size
    part_pred = [(p["category"], 12, 12)] if q == 0 else [(p["brand1"], 2221, 2228)] if q == 1 else \
        [(p["brand1"], 2239, 2239)]
    region = [1, 2, 3][q]
    enc = dict(codec=1, page_rows=4096, chunk_rows=262144, lz4_min_match=4)
    cols = {k: ctx.upload(golap.HostColumn.encode(lo[k], **enc)) for k in ("orderdate", "partkey", "suppkey", "revenue")}
    maps, bufs = [], []
    for key, n_dim, preds, pay in ((d["key"], 2555, [], d["year"]),
                                   (np.arange(n_parts, dtype=np.int32), n_parts, part_pred, p["brand1"]),
                                   (np.arange(n_supp, dtype=np.int32), n_supp, [(s["region"], region, region)], None)):
        m, b = build_map(ctx, golap, key, n_dim, preds, pay)
        maps.append(m)
        bufs += b
    try:
        dims = [(cols["orderdate"], maps[0], 1992, 7), (cols["partkey"], maps[1], 1101, 4500),
                (cols["suppkey"], maps[2], 0, 0)]
        r, k, sums, cnt = ctx.star_run(dims, cols["revenue"])
        assert (r.total_sum, r.n_groups) == SF10_Q2[q]
        # group by group: the oracle's CPU plan over the same tables (pinned by the same log)
        wk, ws, wc = orc.star_groupby([(d["key"], [], d["year"]),
                                       (np.arange(n_parts, dtype=np.int32), part_pred, p["brand1"]),
                                       (np.arange(n_supp, dtype=np.int32), [(s["region"], region, region)], None)],
                                      [lo["orderdate"], lo["partkey"], lo["suppkey"]], lo["revenue"], threads=8)
        assert np.array_equal(k, wk) and np.array_equal(sums, ws) and np.array_equal(cnt, wc)
    finally:
        for c_ in cols.values():
            ctx.free_column(c_)
        for m in maps:
            ctx.free_hashmap(m)
        for b in bufs:
            ctx.free(b)


def _ki(city):
    """CityKIPred (query_common.cuh:73-82): city 'UNITED KI1' or 'UNITED KI5' (221, 225)."""
    return ((city == 221) | (city == 225)).astype(np.uint8)


# SSB Q3.x / Q4.x (03.hpp:611-677, 04.hpp:316-381): per dimension (predicates, group attribute,
# group range); dimensions in group-key order. The "IN (KI1, KI5)" city predicate is a derived
# 0/1 column of the dimension (CityKIPred evaluated over the dimension rows).
def _q34(q, d, p, s, c):
    date = ("date", d["key"])
    cust = ("cust", np.arange(c["region"].size, dtype=np.int32))
    supp = ("supp", np.arange(s["region"].size, dtype=np.int32))
    part = ("part", np.arange(p["category"].size, dtype=np.int32))
    y9297 = [(d["year"], 1992, 1997)]
    y9798 = [(d["year"], 1997, 1998)]
    year = (d["year"], 1992, 7)
    if q == 0:
        return [(cust, [(c["region"], 2, 2)], (c["nation"], 0, 25)), (supp, [(s["region"], 2, 2)], (s["nation"], 0, 25)),
                (date, y9297, year)], False
    if q == 1:
        return [(cust, [(c["nation"], 23, 23)], (c["city"], 0, 250)), (supp, [(s["nation"], 23, 23)], (s["city"], 0, 250)),
                (date, y9297, year)], False
    if q in (2, 3):
        dp = y9297 if q == 2 else [(d["yearmonthnum"], 199712, 199712)]
        return [(cust, [(_ki(c["city"]), 1, 1)], (c["city"], 0, 250)), (supp, [(_ki(s["city"]), 1, 1)], (s["city"], 0, 250)),
                (date, dp, year)], False
    if q == 4:
        return [(date, [], year), (cust, [(c["region"], 1, 1)], (c["nation"], 0, 25)),
                (supp, [(s["region"], 1, 1)], None), (part, [(p["mfgr"], 1, 2)], None)], True
    if q == 5:
        return [(date, y9798, year), (cust, [(c["region"], 1, 1)], None), (supp, [(s["region"], 1, 1)], (s["nation"], 0, 25)),
                (part, [(p["mfgr"], 1, 2)], (p["category"], 11, 45))], True
    return [(date, y9798, year), (cust, [(c["region"], 1, 1)], None), (supp, [(s["nation"], 23, 23)], (s["city"], 0, 250)),
            (part, [(p["category"], 14, 14)], (p["brand1"], 1401, 40))], True


SF10_Q34 = [(7511687354847, 150), (301286847125, 600), (12099576386, 24), (140730476, 4),  # query_sf10_export.csv
            (3440778955099, 35), (981566989074, 100), (19815949200, 798)]


@pytest.mark.parametrize("q", range(7), ids=["q3.1", "q3.2", "q3.3", "q3.4", "q4.1", "q4.2", "q4.3"])
def test_ssb_sf10_q3_q4_reproduce_reference_log(orc, golap, ctx, ssb10, q):
    """pipeline_q3 (03.cuh:118-170): SUM(lo_revenue) over customer, supplier and date; pipeline_q4
    (04.cuh:99-153): SUM(lo_revenue - lo_supplycost) over date, customer, supplier and part, the
    difference taken on the device (the plan's value_minus column). The logged sums and group
    counts, and every group against the oracle's CPU plan over the same tables."""
    lo, d, p, s, c = ssb10
    dims, q4 = _q34(q, d, p, s, c)
    say("ssb sf10", ["q3.1", "q3.2", "q3.3", "q3.4", "q4.1", "q4.2", "q4.3"][q], "rows", lo["orderdate"].size)
    fk_of = dict(date="orderdate", cust="custkey", supp="suppkey", part="partkey")
    enc = dict(codec=1, page_rows=4096, chunk_rows=262144, lz4_min_match=4)
    names = [fk_of[t] for (t, _), _, _ in dims] + ["revenue"] + (["supplycost"] if q4 else [])
    cols = {k: ctx.upload(golap.HostColumn.encode(lo[k], **enc)) for k in names}
    maps, bufs = [], []
    try:
        plan = []
        for (t, key), preds, grp in dims:
            m, b = build_map(ctx, golap, key, key.size, preds, None if grp is None else grp[0])
            maps.append(m)
            bufs += b
            plan.append((cols[fk_of[t]], m, 0 if grp is None else grp[1], 0 if grp is None else grp[2]))
        r, k, sums, cnt = ctx.star_run(plan, cols["revenue"], value_minus=cols["supplycost"] if q4 else None)
        assert (r.total_sum, r.n_groups) == SF10_Q34[q]
        val = lo["revenue"].astype(np.int64) - (lo["supplycost"].astype(np.int64) if q4 else 0)
        assert val.min() >= 0 and val.max() < 2**31
        wk, ws, wc = orc.star_groupby([(key, preds, None if grp is None else grp[0]) for (t, key), preds, grp in dims],
                                      [lo[fk_of[t]] for (t, _), _, _ in dims], val.astype(np.int32), threads=8)
        assert np.array_equal(k, wk) and np.array_equal(sums, ws) and np.array_equal(cnt, wc)
    finally:
        for c_ in cols.values():
            ctx.free_column(c_)
        for m in maps:
            ctx.free_hashmap(m)
        for b in bufs:
            ctx.free(b)


def test_value_minus_refused_outside_single_plans(golap, ctx):
    """value_minus is a golap_star_run option: a batch carrying it is refused, not run without it."""
    rng = np.random.default_rng(5)
    n = 100_000
    fk = rng.integers(0, 100, n).astype(np.int32)
    v = rng.integers(0, 1000, n).astype(np.int32)
    enc = dict(codec=0, page_rows=4096, chunk_rows=65536)
    cf, cv = ctx.upload(golap.HostColumn.encode(fk, **enc)), ctx.upload(golap.HostColumn.encode(v, **enc))
    m, bufs = build_map(ctx, golap, np.arange(100, dtype=np.int32), 100, [], None)
    try:
        r, _, sums, _ = ctx.star_run([(cf, m, 0, 0)], cv, value_minus=cv)
        assert r.total_sum == 0 and r.rows_matched == n
        with pytest.raises(golap.GolapError):
            ctx.star_batch([dict(dims=[(cf, m, 0, 0)], value=cv, value_minus=cv)])
    finally:
        ctx.free_column(cf)
        ctx.free_column(cv)
        ctx.free_hashmap(m)
        for b in bufs:
            ctx.free(b)
