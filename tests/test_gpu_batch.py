"""Batched star plans: many queries over one date-sorted fact table in one persistent launch
(work list built on the device from each query's zone-map predicate, hipGraph replay, autotuned
launch configuration). Every query must equal the single-query plan (golap_star_run with the
same zone-map mask, check_mmmeta semantics of core/metadata.cuh:539-547) and the oracle's
restatement of the CPU plan (02inmem.hpp:58-206) over the query's slice."""
import numpy as np
import pytest

from gpu_util import say
from test_gpu_star import build_map, star_data

pytestmark = pytest.mark.gpu

CHUNK = 65536


def make_queries(rng, nq, nch):
    """Seeded query mix: date range (zone-map predicate and date map), slice, dimension subset."""
    qs = []
    for i in range(nq):
        y0 = int(rng.integers(1992, 1999))
        y1 = min(1998, y0 + int(rng.integers(0, 3)))
        cb = int(rng.integers(0, nch - 1))
        length = int(np.exp(rng.uniform(0, np.log(nch - cb))))
        ce = min(nch, cb + max(1, length))
        kind = i % 3
        qs.append(dict(y0=y0, y1=y1, cb=cb, ce=ce, kind=kind, region=int(rng.integers(0, 5)),
                       nation=int(rng.integers(0, 25)), blo=int(rng.integers(1, 900)), bw=int(rng.integers(0, 100))))
    return qs


class Fixture:
    def __init__(self, golap, ctx, n, seed, date_codec=2, codec=1):
        self.golap, self.ctx = golap, ctx
        self.D = star_data(golap, n, seed=seed)
        enc = dict(page_rows=4096, chunk_rows=CHUNK, lz4_min_match=16)
        self.cols = {nm: ctx.upload(golap.HostColumn.encode(self.D[nm], codec=date_codec if nm == "date" else codec,
                                                            **enc))
                     for nm in ("date", "cust", "supp", "part", "rev")}
        self.nch = ctx.column_desc(self.cols["date"]).n_chunks
        self.maps, self.bufs, self.cache = [], [], {}

    def map(self, name, key, n, preds, pay, tag):
        if tag not in self.cache:
            m, b = build_map(self.ctx, self.golap, key, n, preds, pay)
            self.maps.append(m)
            self.bufs += b
            self.cache[tag] = m
        return self.cache[tag]

    def spec(self, q):
        """(dims for the library, dims for the oracle, fk names) of one query."""
        D = self.D
        dmap = self.map("date", D["d_key"], 2556, [(D["d_year"], q["y0"], q["y1"])], D["d_year"],
                        ("d", q["y0"], q["y1"]))
        dims = [(self.cols["date"], dmap, 1992, 7)]
        odims = [(D["d_key"], [(D["d_year"], q["y0"], q["y1"])], D["d_year"])]
        names = ["date"]
        if q["kind"] in (0, 2):
            smap = self.map("supp", D["s_key"], D["s_key"].size, [(D["s_region"], q["region"], q["region"])], None,
                            ("s", q["region"]))
            dims.append((self.cols["supp"], smap, 0, 0))
            odims.append((D["s_key"], [(D["s_region"], q["region"], q["region"])], None))
            names.append("supp")
        if q["kind"] in (1, 2):
            cmap = self.map("cust", D["c_key"], D["c_key"].size, [(D["c_nation"], q["nation"], q["nation"])], None,
                            ("c", q["nation"]))
            dims.append((self.cols["cust"], cmap, 0, 0))
            odims.append((D["c_key"], [(D["c_nation"], q["nation"], q["nation"])], None))
            names.append("cust")
        if q["kind"] in (0, 2):
            blo, bhi = q["blo"], q["blo"] + q["bw"]
            pmap = self.map("part", D["p_key"], D["p_key"].size, [(D["p_brand"], blo, bhi)], D["p_brand"],
                            ("p", blo, bhi))
            dims.append((self.cols["part"], pmap, 1, 1000))
            odims.append((D["p_key"], [(D["p_brand"], blo, bhi)], D["p_brand"]))
            names.append("part")
        return dims, odims, names

    def batch_entry(self, q):
        dims, _, _ = self.spec(q)
        lo, hi = q["y0"] * 10000 + 101, q["y1"] * 10000 + 1231
        return dict(dims=dims, value=self.cols["rev"], chunk_begin=q["cb"], chunk_end=q["ce"],
                    zone=(self.cols["date"], lo, hi))

    def single(self, q, flags):
        dims, _, _ = self.spec(q)
        lo, hi = q["y0"] * 10000 + 101, q["y1"] * 10000 + 1231
        self.ctx.minmax_check(self.cols["date"], lo, hi, flags)
        return self.ctx.star_run(dims, self.cols["rev"], chunk_begin=q["cb"], chunk_end=q["ce"], dev_chunk_mask=flags)

    def oracle(self, orc, q):
        _, odims, names = self.spec(q)
        r0, r1 = q["cb"] * CHUNK, min(self.D["date"].size, q["ce"] * CHUNK)
        return orc.star_groupby(odims, [self.D[nm][r0:r1] for nm in names], self.D["rev"][r0:r1], threads=4)

    def free(self):
        for c in self.cols.values():
            self.ctx.free_column(c)
        for m in self.maps:
            self.ctx.free_hashmap(m)
        for b in self.bufs:
            self.ctx.free(b)


@pytest.fixture(scope="module")
def fx(golap, ctx):
    f = Fixture(golap, ctx, 3_000_000, seed=7)
    yield f
    f.free()


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a[1:], b[1:]))


def test_batch_matches_single_and_oracle(orc, golap, ctx, fx):
    qs = make_queries(np.random.default_rng(11), 12, fx.nch)
    say("batch of", len(qs), "queries over", fx.nch, "chunks")
    flags = ctx.alloc(fx.nch * 2)
    try:
        with ctx.star_batch([fx.batch_entry(q) for q in qs]) as b:
            info = b.run()
            for i, q in enumerate(qs):
                got = b.result(i)
                ref = fx.single(q, flags)
                assert same(got, ref), i
                assert got[0].rows_matched == ref[0].rows_matched and got[0].total_sum == ref[0].total_sum
                assert (got[0].chunks_run, got[0].chunks_pruned, got[0].comp_bytes) == \
                       (ref[0].chunks_run, ref[0].chunks_pruned, ref[0].comp_bytes), i
                if i < 4:
                    wk, ws, wc = fx.oracle(orc, q)
                    assert np.array_equal(got[1], wk) and np.array_equal(got[2], ws) and np.array_equal(got[3], wc)
            # the device work list holds the surviving chunks' parts: one item per (query,
            # chunk part)
            cfg = b.config()
            assert cfg.shared_filter == 0
            parts = -(-16 // cfg.pages_per_item)
            runs = [b.result(i)[0].chunks_run for i in range(len(qs))]
            assert info.items == sum(runs) * parts
            assert info.comp_bytes == sum(b.result(i)[0].comp_bytes for i in range(len(qs)))
            # with the shared filter (every query probes the date column): one item per chunk
            # part that any query keeps, and the same groups
            b.set_config(shared_filter=1)
            info = b.run()
            assert info.items % parts == 0 and max(runs) * parts <= info.items < sum(runs) * parts
            for i, q in enumerate(qs):
                assert same(b.result(i), fx.single(q, flags)), ("shared filter", i)
    finally:
        ctx.free(flags)


def test_batch_graph_replays_and_plain_launches_agree(golap, ctx, fx):
    qs = make_queries(np.random.default_rng(12), 9, fx.nch)
    flags = ctx.alloc(fx.nch * 2)
    try:
        with ctx.star_batch([fx.batch_entry(q) for q in qs]) as b:
            b.run(use_graph=False)
            want = [b.result(i) for i in range(len(qs))]
            for rep in range(3):
                say("graph replay", rep)
                b.run(use_graph=True)
                for i in range(len(qs)):
                    assert same(b.result(i), want[i]), (rep, i)
                # every query's groups from one call equal the per-query ones, also when the
                # first buffer is too small
                for got, w in zip(b.results(), want):
                    assert same(got, w) and got[0].n_groups == w[0].n_groups
                for got, w in zip(b.results(max_total=3), want):
                    assert same(got, w)
                fx.single(qs[rep], flags)  # eager work of the program between replays
    finally:
        ctx.free(flags)


CONFIGS = [
    dict(pages_per_item=1, lds_group_slots=0, filter_lds=0, compact_tables=0, shared_filter=0),
    dict(pages_per_item=4, lds_group_slots=1024, lds_cutover=0, filter_lds=1, compact_tables=1, shared_filter=0),
    dict(pages_per_item=16, lds_group_slots=4096, lds_cutover=0, max_blocks_per_cu=1, shared_filter=0),
    dict(pages_per_item=2, lds_group_slots=2048, lds_cutover=1 << 30, filter_lds=0, max_blocks_per_cu=2,
         shared_filter=0),
    # one membership test per row for all queries (every query here probes the date dimension)
    dict(pages_per_item=1, compact_tables=0, shared_filter=1),
    dict(pages_per_item=16, compact_tables=1, max_blocks_per_cu=2, shared_filter=1),
]


def test_batch_configurations_agree(golap, ctx, fx):
    qs = make_queries(np.random.default_rng(13), 10, fx.nch)
    with ctx.star_batch([fx.batch_entry(q) for q in qs]) as b:
        b.run()
        want = [b.result(i) for i in range(len(qs))]
        for cfg in CONFIGS:
            say("config", cfg)
            b.set_config(**cfg)
            b.run()
            for i in range(len(qs)):
                assert same(b.result(i), want[i]), (cfg, i)
        c, ms = b.tune(reps=2)
        say("tuned", c.as_dict(), "ms", ms)
        b.run()
        for i in range(len(qs)):
            assert same(b.result(i), want[i]), i
        # the tuned configuration is cached for the same batch shape
        with ctx.star_batch([fx.batch_entry(q) for q in qs]) as b2:
            assert b2.config().as_dict() == c.as_dict()


def test_batch_rejects_bad_config_and_rebuilt_map(golap, ctx, fx):
    qs = make_queries(np.random.default_rng(14), 3, fx.nch)
    with ctx.star_batch([fx.batch_entry(q) for q in qs]) as b:
        before = b.config().as_dict()
        with pytest.raises(golap.GolapError):
            b.set_config(pages_per_item=0)
        with pytest.raises(golap.GolapError):
            b.set_config(lds_group_slots=3000)
        assert b.config().as_dict() == before
        b.run()
        # rebuilding a map the batch probes invalidates the batch
        dims = fx.batch_entry(qs[0])["dims"]
        m = dims[0][1]
        dy = ctx.to_device(fx.D["d_year"])
        try:
            # same predicate and payload: same contents, but a new build
            ctx.hashmap_build(m, 2556, [(dy, golap.I32, qs[0]["y0"], qs[0]["y1"])], dy)
            with pytest.raises(golap.GolapError):
                b.run()
        finally:
            ctx.free(dy)


def test_batch_many_queries_one_chunk_slices(golap, ctx, fx):
    """64 queries, many of one chunk, some with empty zone ranges (everything pruned)."""
    rng = np.random.default_rng(15)
    qs = make_queries(rng, 64, fx.nch)
    for i in range(0, 64, 5):
        qs[i]["ce"] = qs[i]["cb"] + 1
    entries = [fx.batch_entry(q) for q in qs]
    for i in range(0, 64, 7):
        entries[i]["zone"] = (fx.cols["date"], 30000101, 30001231)  # no chunk survives
    flags = ctx.alloc(fx.nch * 2)
    try:
        with ctx.star_batch(entries) as b:
            b.run()
            for i, q in enumerate(qs):
                got = b.result(i)
                if i % 7 == 0:
                    assert got[0].n_groups == 0 and got[0].chunks_run == 0
                    continue
                assert same(got, fx.single(q, flags)), i
    finally:
        ctx.free(flags)
