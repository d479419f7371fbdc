"""Prejoined metadata (SortHelper::prejoin_tables, usecases/ssb/helper.hpp:115-152, with
MinMaxMeta on the prejoined column, MetaFlags::META, usecases/ssb/queries/01.hpp:70-78,146-150):
per-chunk ranges of a dimension attribute reached through the join, and the pruning they allow
on a fact table sorted by prejoined attributes (SortHelper::apply "general_dimsort",
helper.hpp:225-233)."""
import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu


def _year_map(ctx, golap, key, year, preds=()):
    dk = ctx.to_device(key)
    bufs = [dk]
    dp = []
    for col, lo, hi in preds:
        d = ctx.to_device(col)
        bufs.append(d)
        dp.append((d, golap.TYPE_OF[col.dtype], lo, hi))
    pay = ctx.to_device(year.astype(np.int32))
    bufs.append(pay)
    m = ctx.hashmap(key.size, dk, golap.I32)
    ctx.hashmap_build(m, key.size, dp, pay)
    return m, bufs


def test_prejoin_minmax_matches_numpy(golap, ctx):
    """Per chunk, min and max of year[fk] over the rows whose key the map holds (a predicate
    leaves some keys out; a chunk of missing keys has no range and never qualifies)."""
    r = np.random.default_rng(91)
    key = (np.arange(3000, dtype=np.int32) * 7 + 11)
    year = (1990 + r.integers(0, 30, key.size)).astype(np.int16)
    keep = r.integers(0, 4, key.size).astype(np.uint8)  # keys with keep == 0 are not in the map
    n, chunk = 1_000_003, 65536
    idx = np.sort(r.integers(0, key.size, n))  # sorted: chunks cover narrow key ranges
    idx[5 * chunk:6 * chunk] = np.flatnonzero(keep == 0)[0]  # a chunk whose rows all miss
    fk = key[idx]
    say("prejoin minmax rows", n)
    col = ctx.upload(golap.HostColumn.encode(fk, codec=1, page_rows=4096, chunk_rows=chunk, lz4_min_match=4))
    m, bufs = _year_map(ctx, golap, key, year, [(keep, 1, 3)])
    h = ctx.prejoin_minmax(col, m)
    flags = ctx.alloc(2 * ((n + chunk - 1) // chunk))
    try:
        d = ctx.chunkmeta_desc(h)
        nc = d.n_chunks
        assert nc == (n + chunk - 1) // chunk
        mins = ctx.to_host(d.mins, nc, np.int64)
        maxs = ctx.to_host(d.maxs, nc, np.int64)
        hit = keep[idx] != 0
        for c in range(nc):
            sl = slice(c * chunk, min(n, (c + 1) * chunk))
            ys = year[idx[sl]][hit[sl]].astype(np.int64)
            if ys.size:
                assert (mins[c], maxs[c]) == (ys.min(), ys.max())
            else:
                assert mins[c] > maxs[c]
        for lo, hi in ((2000, 2005), (1990, 1990), (2019, 2030), (1900, 1989)):
            ctx.chunkmeta_minmax_check(h, lo, hi, flags)
            got = ctx.to_host(flags, nc, np.uint16)
            assert np.array_equal(got, ((lo <= maxs) & (hi >= mins)).astype(np.uint16))
        assert got.sum() == 0
    finally:
        ctx.free(flags)
        ctx.chunkmeta_free(h)
        ctx.free_hashmap(m)
        for b in bufs:
            ctx.free(b)
        ctx.free_column(col)


@pytest.fixture(scope="module")
def ssb10_dimsort(orc):
    """SSB SF10 with the fact table sorted like SortHelper::apply("general_dimsort"): by the
    prejoined (c_region, lo_discount, c_nation, lo_quantity, d_yearmonthnum, c_city)."""
    lo, d, p, s, c = orc.ssb_generate(10)
    di = np.searchsorted(d["key"], lo["orderdate"])
    cust = lo["custkey"]
    order = np.lexsort((c["city"][cust], d["yearmonthnum"][di], lo["quantity"], c["nation"][cust], lo["discount"],
                        c["region"][cust]))
    return {k: np.ascontiguousarray(v[order]) for k, v in lo.items()}, d


def test_ssb_sf10_q1_prejoined_pruning(orc, golap, ctx, ssb10_dimsort):
    """Q1.1 on the dimension-sorted fact table in chunks of 131,072 rows (1 MiB of the widest
    column at the reference's SF200 run): d_year = 1993 through the prejoined year's per-chunk
    range, lo_discount 1..3 and lo_quantity 0..24 through the zone maps; the pruned plan gives the
    logged sum, as the unpruned one does, and the skipped share of compressed bytes is reported
    (the reference logs 86.4% at SF200, usecases/ssb/results/pruning_query1123_export.csv)."""
    lo, d = ssb10_dimsort
    n = lo["orderdate"].size
    chunk = 131072
    enc = dict(codec=1, page_rows=4096, chunk_rows=chunk, lz4_min_match=4)
    cols = [ctx.upload(golap.HostColumn.encode(lo[k], **enc)) for k in ("orderdate", "discount", "quantity",
                                                                       "extendedprice")]
    pm, pbufs = _year_map(ctx, golap, d["key"], d["year"])  # the prejoin: payload d_year
    qm, qbufs = _year_map(ctx, golap, d["key"], d["year"], [(d["year"], 1993, 1993)])
    nc = ctx.column_desc(cols[0]).n_chunks
    f1, f2 = ctx.alloc(2 * nc), ctx.alloc(2 * nc)
    h = ctx.prejoin_minmax(cols[0], pm)
    try:
        ctx.chunkmeta_minmax_check(h, 1993, 1993, f1)
        ctx.minmax_check(cols[1], 1, 3, f2)
        ctx.combine_and(f1, f2, nc)
        ctx.minmax_check(cols[2], 0, 24, f2)
        ctx.combine_and(f1, f2, nc)
        keep = ctx.to_host(f1, nc, np.uint16)
        full = ctx.q1_run(*cols, qm, (1, 3), (0, 24))
        pruned = ctx.q1_run(*cols, qm, (1, 3), (0, 24), dev_chunk_mask=f1)
        assert full.sum == pruned.sum == 4212228344835  # query_sf10_export.csv, Q1.1
        assert pruned.chunks_pruned == nc - int(keep.sum())
        skipped = 1 - pruned.comp_bytes / full.comp_bytes
        say(f"sf10 q1.1 dimsort: {nc} chunks, {pruned.chunks_pruned} pruned, {skipped:.1%} of compressed bytes "
            f"skipped (reference SF200: 86.4%)")
        assert 0.5 < skipped < 1
        # adding the prejoined range only removes chunks: every kept chunk passes the discount map
        ctx.minmax_check(cols[1], 1, 3, f2)
        nf = ctx.to_host(f2, nc, np.uint16)
        assert np.all(nf >= keep)
    finally:
        ctx.chunkmeta_free(h)
        for b in (f1, f2):
            ctx.free(b)
        for m_ in (pm, qm):
            ctx.free_hashmap(m_)
        for b in pbufs + qbufs:
            ctx.free(b)
        for c_ in cols:
            ctx.free_column(c_)
