"""Fused scan + filter + group-by over the trip table (range predicates of 01.cuh:26-54,
grouping of group_by_agg_2col/4col + HashAggregate::add, core/group.cuh:14-36 and
core/dev_structs.cuh:209-244, over computed keys: zone x hour of the pickup timestamp) against
the oracle's restatement."""
import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu

T0 = 1704067200


def trip_cols(golap, ctx, n, seed, codec=255, page_rows=4096, chunk_rows=65536):
    host = golap.gen_trip(seed, n)
    dev = [ctx.upload(golap.HostColumn.encode(c, codec=codec, page_rows=page_rows, chunk_rows=chunk_rows))
           for c in host]
    return host, dev


def template(t, h0, h1, full_hours=False):
    """(preds, keys, value_col) of the four request templates over hours [h0, h1] of the year.
    Columns: 0 pickup_ts, 1 passenger, 2 zone, 3 fare, 4 distance."""
    lo, hi = T0 + h0 * 3600, T0 + (h1 + 1) * 3600 - 1
    hb, hc = (0, 8784) if full_hours else (h0, h1 - h0 + 1)
    zone = (2, 1, 1, 265, 1)
    hour = (0, T0 + hb * 3600, 3600, hc, hb)
    if t == 0:
        return [(0, lo, hi)], [zone, hour], 3
    if t == 1:
        return [(0, lo, hi), (1, 2, 6)], [zone, hour], 4
    if t == 2:
        return [(0, lo, hi), (3, 1000, 5000)], [hour, zone], 4
    return [(0, lo, hi), (4, 5000, 1 << 30)], [(1, 1, 1, 6, 1), zone], -1


def check(orc, ctx, host, dev, preds, keys, vc, **kw):
    r, k, s, c = ctx.scan_run(dev, preds, keys, vc, **kw)
    cb, ce = kw.get("chunk_begin", 0), kw.get("chunk_end")
    rows = slice(cb * 65536, None if ce is None else ce * 65536)
    w = orc.scan_groupby([h[rows] for h in host], preds, keys, vc, threads=4)
    assert w is not None
    wk, ws, wc = w
    assert np.array_equal(k, wk) and np.array_equal(s, ws) and np.array_equal(c, wc)
    assert r.rows_matched == int(wc.sum()) and r.total_sum == int(ws.astype(np.uint64).sum().astype(np.int64))
    return r


@pytest.fixture(scope="module")
def trip(golap, ctx):
    host, dev = trip_cols(golap, ctx, 3_000_000, seed=5)
    yield host, dev
    for d in dev:
        ctx.free_column(d)


@pytest.mark.parametrize("t", range(4))
@pytest.mark.parametrize("agg", [0, 2])
def test_scan_templates_match_oracle(orc, golap, ctx, trip, t, agg):
    host, dev = trip
    # 3M rows span the year: hours 1000..1700 hold about 240k rows
    preds, keys, vc = template(t, 1000, 1700)
    say("scan template", t, "agg", agg)
    r = check(orc, ctx, host, dev, preds, keys, vc, agg_mode=agg)
    assert r.chunks_pruned > r.chunks_run  # the sorted timestamp prunes most zones


def test_scan_full_year_groups_spill(orc, golap, ctx, trip):
    """Zone x hour over the whole year: 2,327,760 groups, LDS tables flushed and spilled."""
    host, dev = trip
    preds, keys, vc = template(0, 0, 8783, full_hours=True)
    r = check(orc, ctx, host, dev, preds, keys, vc)
    assert r.lds_group_slots > 0 and r.n_groups > 1_000_000


@pytest.mark.parametrize("slots", [64, 512, 4096])
def test_scan_small_lds_tables_agree(orc, golap, ctx, trip, slots):
    host, dev = trip
    preds, keys, vc = template(1, 200, 5000)
    check(orc, ctx, host, dev, preds, keys, vc, agg_mode=1, lds_group_slots=slots)


def test_scan_slice_mask_and_no_prune_agree(orc, golap, ctx, trip):
    host, dev = trip
    preds, keys, vc = template(2, 3000, 3300)
    a = check(orc, ctx, host, dev, preds, keys, vc, chunk_begin=3, chunk_end=40)
    b = check(orc, ctx, host, dev, preds, keys, vc, chunk_begin=3, chunk_end=40, zone_prune=0)
    assert a.chunks_run < b.chunks_run and b.chunks_pruned == 0
    nch = ctx.column_desc(dev[0]).n_chunks
    flags = ctx.alloc(nch * 2)
    try:
        ctx.minmax_check(dev[0], preds[0][1], preds[0][2], flags)
        c = check(orc, ctx, host, dev, preds, keys, vc, chunk_begin=3, chunk_end=40, zone_prune=0,
                  dev_chunk_mask=flags)
        assert c.chunks_run == a.chunks_run
    finally:
        ctx.free(flags)


@pytest.mark.parametrize("codec", [0, 1, 2, 4])
def test_scan_codecs(orc, golap, ctx, codec):
    host, dev = trip_cols(golap, ctx, 400_000, seed=9, codec=codec, page_rows=2048, chunk_rows=16384)
    try:
        for t in range(4):
            preds, keys, vc = template(t, 100, 8000)
            r, k, s, c = ctx.scan_run(dev, preds, keys, vc)
            wk, ws, wc = orc.scan_groupby(list(host), preds, keys, vc, threads=4)
            assert np.array_equal(k, wk) and np.array_equal(s, ws) and np.array_equal(c, wc), (codec, t)
    finally:
        for d in dev:
            ctx.free_column(d)


@pytest.mark.parametrize("page_rows", [1024, 4096])
def test_scan_lz4i_pages_decode_in_parallel(orc, golap, ctx, page_rows):
    """Every column LZ4-coded at minimum match 4 (LZ4I pages: timestamps with thousands of
    sequences a page, passenger counts with tens): the scan's windows stage and decode them in
    parallel, single plans and a batch of them, against the oracle."""
    host, dev = trip_cols(golap, ctx, 700_000, seed=19, codec=1, page_rows=page_rows, chunk_rows=65536)
    try:
        codes = [ctx.to_host(ctx.column_desc(d).page_codec, ctx.column_desc(d).n_pages, np.uint8) for d in dev]
        assert all((c == 5).mean() > 0.9 for c in codes[:2])  # ts and passenger pages carry an index
        plans = []
        for t in range(4):
            preds, keys, vc = template(t, 500, 7000)
            say("scan lz4i template", t, "page_rows", page_rows)
            check(orc, ctx, host, dev, preds, keys, vc)
            plans.append(dict(cols=dev, preds=preds, keys=keys, value_col=vc))
        with ctx.scan_batch(plans) as b:
            b.run()
            for i, p in enumerate(plans):
                got, ref = b.result(i), ctx.scan_run(**p)
                assert all(np.array_equal(x, y) for x, y in zip(got[1:], ref[1:])), i
    finally:
        for d in dev:
            ctx.free_column(d)


def test_scan_batch_reports_corrupt_page_of_a_decoded_column(orc, golap, ctx):
    """A scan batch of several plans decodes its LZ4I columns ahead of the launch; a truncated
    passenger page there is reported as ECORRUPT by the run, and the batch over intact columns
    still equals the single-plan runs afterwards."""
    host = golap.gen_trip(23, 300_000)
    good, bad = [], []
    try:
        for i, c in enumerate(host):
            hc = golap.HostColumn.encode(c, codec=1 if i == 1 else 255, page_rows=4096, chunk_rows=65536)
            good.append(ctx.upload(hc))
            if i == 1:
                a = hc.arrays()
                n = a["desc"].n_pages
                pages = [bytes(a["comp"][int(a["page_off"][p]):int(a["page_off"][p]) + int(a["page_csize"][p])])
                         for p in range(n)]
                assert int(a["page_codec"][5]) == 5  # an LZ4I page
                pages[5] = pages[5][:len(pages[5]) // 2]
                hc = golap.HostColumn.assemble(golap.U8, c.size, 4096, 65536, pages, a["page_codec"].tolist())
                bad.append(ctx.upload(hc))
            else:
                bad.append(good[-1])
        say("scan batch corrupt decoded page")
        for cols, corrupt in ((bad, True), (good, False)):
            plans = []
            for t in (1, 3):
                preds, keys, vc = template(t, 0, 8783, full_hours=True)
                plans.append(dict(cols=cols, preds=preds, keys=keys, value_col=vc))
            with ctx.scan_batch(plans) as b:
                if corrupt:
                    with pytest.raises(golap.GolapError) as e:
                        b.run()
                    assert e.value.status == golap.ECORRUPT
                else:
                    b.run()
                    for i, p in enumerate(plans):
                        got, ref = b.result(i), ctx.scan_run(**p)
                        assert all(np.array_equal(x, y) for x, y in zip(got[1:], ref[1:])), i
    finally:
        ctx.free_column(bad[1])
        for d in good:
            ctx.free_column(d)


def test_scan_key_out_of_range_is_an_error(golap, ctx, trip):
    host, dev = trip
    preds, keys, vc = template(0, 1000, 1700)
    keys = [keys[0], (0, T0 + 1000 * 3600, 3600, 100, 1000)]  # hours 1100..1700 fall outside
    with pytest.raises(golap.GolapError):
        ctx.scan_run(dev, preds, keys, vc)
    r = ctx.scan_run(dev, *template(0, 1000, 1700))  # the context still works
    assert r[0].n_groups > 0


def scan_mix(rng, n):
    """Seeded trip requests: template 0..3 over an hour range of log-uniform length."""
    out = []
    for i in range(n):
        t = int(rng.integers(0, 4))
        span = int(np.exp(rng.uniform(0, np.log(8784))))
        h0 = int(rng.integers(0, 8784 - span + 1))
        out.append((t, h0, h0 + span - 1, bool(i % 5 == 0)))
    return out


def test_scan_batch_matches_single_and_oracle(orc, golap, ctx, trip):
    host, dev = trip
    mix = scan_mix(np.random.default_rng(21), 16)
    plans = []
    for t, h0, h1, full in mix:
        preds, keys, vc = template(t, h0, h1, full_hours=full)
        plans.append(dict(cols=dev, preds=preds, keys=keys, value_col=vc))
    with ctx.scan_batch(plans) as b:
        b.run()
        for i, p in enumerate(plans):
            got = b.result(i)
            ref = ctx.scan_run(**p)
            assert all(np.array_equal(x, y) for x, y in zip(got[1:], ref[1:])), i
            assert (got[0].chunks_run, got[0].comp_bytes, got[0].rows_matched) == \
                   (ref[0].chunks_run, ref[0].comp_bytes, ref[0].rows_matched), i
            if i < 3:
                wk, ws, wc = orc.scan_groupby(list(host), p["preds"], p["keys"], p["value_col"], threads=4)
                assert np.array_equal(got[1], wk) and np.array_equal(got[2], ws) and np.array_equal(got[3], wc)
        want = [b.result(i) for i in range(len(plans))]
        # all queries' groups from one call (compacted on the device when sparse) equal them
        for got, w in zip(b.results(), want):
            assert all(np.array_equal(x, y) for x, y in zip(got[1:], w[1:]))
            assert (got[0].n_groups, got[0].rows_matched, got[0].total_sum) == \
                   (w[0].n_groups, w[0].rows_matched, w[0].total_sum)
        for cfg in [dict(lds_group_slots=0), dict(lds_group_slots=256, pages_per_item=1),
                    dict(lds_group_slots=4096, max_blocks_per_cu=1, pages_per_item=4),
                    dict(lds_group_slots=4096, rows_per_lane=8), dict(rows_per_lane=4)]:
            say("scan batch config", cfg)
            b.set_config(**cfg)
            for rep in range(2):
                b.run(use_graph=rep == 1)
                for i in range(len(plans)):
                    got = b.result(i)
                    assert all(np.array_equal(x, y) for x, y in zip(got[1:], want[i][1:])), (cfg, rep, i)
        with pytest.raises(golap.GolapError):
            b.set_config(rows_per_lane=5)  # rows per lane are 4 or 8
        c, ms = b.tune(reps=2)
        say("scan batch tuned", c.as_dict(), ms)
        assert c.rows_per_lane in (0, 4, 8)  # the tuner's rows-per-wave axis (0: the launch-size rule kept)
        b.run()
        for i in range(len(plans)):
            got = b.result(i)
            assert all(np.array_equal(x, y) for x, y in zip(got[1:], want[i][1:])), i


def test_scan_batch_mixed_column_sets(orc, golap, ctx, trip):
    """Plans that list different subsets of the table's columns share one launch."""
    host, dev = trip
    ts, pas, zone, fare, dist = dev
    plans = [dict(cols=[ts, zone, fare], preds=[(0, T0, T0 + 86400 * 30)], keys=[(1, 1, 1, 265, 1)], value_col=2),
             dict(cols=[dist, ts], preds=[(1, T0 + 86400 * 100, T0 + 86400 * 200), (0, 1000, 9000)],
                  keys=[(1, T0 + 86400 * 100, 86400, 101, 100)], value_col=0),
             dict(cols=[pas, zone], preds=[(0, 3, 6)], keys=[(0, 1, 1, 6, 1), (1, 1, 1, 265, 1)], value_col=-1)]
    with ctx.scan_batch(plans) as b:
        b.run()
        for i, p in enumerate(plans):
            got = b.result(i)
            ref = ctx.scan_run(**p)
            assert all(np.array_equal(x, y) for x, y in zip(got[1:], ref[1:])), i
            hosts = [host[dev.index(c)] for c in p["cols"]]
            wk, ws, wc = orc.scan_groupby(hosts, p["preds"], p["keys"], p["value_col"], threads=4)
            assert np.array_equal(got[1], wk) and np.array_equal(got[2], ws) and np.array_equal(got[3], wc)
        # small domains: the bulk call copies the dense results
        for i, got in enumerate(b.results()):
            assert all(np.array_equal(x, y) for x, y in zip(got[1:], b.result(i)[1:])), i


def test_scan_long_launch_four_row_form(orc, golap, ctx):
    """A table with more pages than 16 per CU: long launches take the 4-rows-per-lane kernel
    (single plans, and a batch whose dense entries leave room for four workgroups per CU)."""
    cus, _ = ctx.info()
    n = (cus * 16 + 256) * 4096  # pages beyond one per wave of four workgroups per CU
    codecs = [golap.CODEC_DBP, golap.CODEC_LZ4, golap.CODEC_FOR, golap.CODEC_FOR, golap.CODEC_FOR]
    host = golap.gen_trip(77, n)
    dev = [ctx.upload(golap.HostColumn.encode(c, codec=k, page_rows=4096, chunk_rows=65536, lz4_min_match=16))
           for c, k in zip(host, codecs)]
    try:
        plans = []
        for t in range(4):
            preds, keys, vc = template(t, 0, 8783, full_hours=True)
            say("long scan template", t, "rows", n)
            r, k, s, c = ctx.scan_run(dev, preds, keys, vc)
            wk, ws, wc = orc.scan_groupby(list(host), preds, keys, vc, threads=8)
            assert np.array_equal(k, wk) and np.array_equal(s, ws) and np.array_equal(c, wc), t
            plans.append(dict(cols=dev, preds=preds, keys=keys, value_col=vc))
        # mixed plans in one launch with 1024 dense entries (four workgroups per CU)
        for t in range(4):
            preds, keys, vc = template(t, 1000 * t, 1000 * t + 3000)
            plans.append(dict(cols=dev, preds=preds, keys=keys, value_col=vc))
        with ctx.scan_batch(plans) as b:
            b.set_config(lds_group_slots=1024)
            b.run()
            for i, p in enumerate(plans):
                got, ref = b.result(i), ctx.scan_run(**p)
                assert all(np.array_equal(x, y) for x, y in zip(got[1:], ref[1:])), i
    finally:
        for d in dev:
            ctx.free_column(d)
