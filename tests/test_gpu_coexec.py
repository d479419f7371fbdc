"""CPU/GPU co-execution (the query 3 variants that keep the customer table on the CPU,
usecases/ssb/queries/03a.hpp, 03c.hpp, 03d.hpp, with HostHashMap and HostHashAggregate,
core/host_structs.hpp:9-152): one dimension joined on the host, the others on the device. The
groups equal the all-device star plan's and the oracle's."""
import numpy as np
import pytest

from gpu_util import say
from test_gpu_star import build_map, star_data

pytestmark = pytest.mark.gpu


def setup(ctx, golap, D, codec=1, chunk_rows=65536, reg=(1, 1), nat=(7, 7), brange=(100, 199)):
    enc = dict(page_rows=4096, chunk_rows=chunk_rows, lz4_min_match=16)
    cols = {name: ctx.upload(golap.HostColumn.encode(D[name], codec=255 if name == "date" else codec, **enc))
            for name in ("date", "cust", "supp", "part", "rev")}
    maps, bufs = [], []
    for key, n_dim, preds, pay in (
            (D["d_key"], 2556, [], D["d_year"]),
            (D["c_key"], D["c_key"].size, [(D["c_nation"], nat[0], nat[1])], None),
            (D["s_key"], D["s_key"].size, [(D["s_region"], reg[0], reg[1])], None),
            (D["p_key"], D["p_key"].size, [(D["p_brand"], brange[0], brange[1])], D["p_brand"])):
        m, b = build_map(ctx, golap, key, n_dim, preds, pay)
        maps.append(m)
        bufs += b
    dims = [(cols["date"], maps[0], 1992, 7), (cols["cust"], maps[1], 0, 0), (cols["supp"], maps[2], 0, 0),
            (cols["part"], maps[3], 1, 1000)]
    host_specs = [(D["d_key"], [], D["d_year"]), (D["c_key"], [(D["c_nation"], nat[0], nat[1])], None),
                  (D["s_key"], [(D["s_region"], reg[0], reg[1])], None),
                  (D["p_key"], [(D["p_brand"], brange[0], brange[1])], D["p_brand"])]
    return cols, maps, bufs, dims, host_specs


def release(ctx, cols, maps, bufs):
    for c in cols.values():
        ctx.free_column(c)
    for m in maps:
        ctx.free_hashmap(m)
    for b in bufs:
        ctx.free(b)


@pytest.mark.parametrize("host_dim", [1, 3, 0])
@pytest.mark.parametrize("threads", [1, 0])
def test_hostdim_equals_device_plan(orc, golap, ctx, host_dim, threads):
    D = star_data(golap, 3_000_000, seed=23)
    cols, maps, bufs, dims, specs = setup(ctx, golap, D, chunk_rows=65536)
    key, preds, pay = specs[host_dim]
    hm = golap.HostHashMap(key)
    hm.build(preds, payload=pay)
    try:
        want = ctx.star_run(dims, cols["rev"])
        got = ctx.star_run_hostdim(dims, cols["rev"], host_dim, hm, threads=threads)
        say("coexec host_dim", host_dim, "threads", threads, got[4].as_dict())
        for a, b in zip(got[1:4], want[1:4]):
            assert np.array_equal(a, b)
        assert (got[0].rows_matched, got[0].total_sum, got[0].n_groups) == \
            (want[0].rows_matched, want[0].total_sum, want[0].n_groups)
        assert got[0].chunks_run == want[0].chunks_run and got[0].comp_bytes == want[0].comp_bytes
        assert got[0].driver_dim != host_dim
        wk, ws, wc = orc.star_groupby(specs, [D["date"], D["cust"], D["supp"], D["part"]], D["rev"], threads=4)
        assert np.array_equal(got[1], wk) and np.array_equal(got[2], ws) and np.array_equal(got[3], wc)
        st = got[4]
        assert st.rounds == 1 and st.emitted >= got[0].rows_matched
    finally:
        hm.free()
        release(ctx, cols, maps, bufs)


@pytest.mark.parametrize("selective", [True, False])
def test_hostdim_rounds_slices_and_masks(orc, golap, ctx, selective):
    """A chunk slice and a chunk mask; with every predicate passing, 20M rows hand over more
    than one round's 4M-row buffer, so rounds are halved until they fit."""
    D = star_data(golap, 20_000_000, seed=31, n_cust=200_000)
    pick = dict(reg=(2, 2), nat=(4, 4), brange=(1, 1000)) if selective else \
        dict(reg=(0, 4), nat=(0, 24), brange=(1, 1000))
    cols, maps, bufs, dims, specs = setup(ctx, golap, D, chunk_rows=262144, **pick)
    key, preds, pay = specs[1]
    hm = golap.HostHashMap(key)
    hm.build(preds, payload=pay)
    nch = ctx.column_desc(cols["rev"]).n_chunks
    mask = (np.arange(nch) % 3 != 1).astype(np.uint16)
    dmask = ctx.to_device(mask)
    try:
        for kw in (dict(), dict(chunk_begin=5, chunk_end=nch - 7), dict(dev_chunk_mask=dmask),
                   dict(chunk_begin=3, chunk_end=40, dev_chunk_mask=dmask)):
            want = ctx.star_run(dims, cols["rev"], **kw)
            got = ctx.star_run_hostdim(dims, cols["rev"], 1, hm, **kw)
            say("coexec rounds", kw.keys(), got[4].as_dict())
            for a, b in zip(got[1:4], want[1:4]):
                assert np.array_equal(a, b)
            assert (got[0].chunks_run, got[0].chunks_pruned) == (want[0].chunks_run, want[0].chunks_pruned)
            assert got[4].rounds >= 1
        full = ctx.star_run_hostdim(dims, cols["rev"], 1, hm)
        if selective:
            assert full[4].rounds == 1 and full[4].splits == 0
        else:
            assert full[4].emitted == full[0].rows_matched == 20_000_000
            assert full[4].splits >= 3 and full[4].rounds == full[4].splits + 1
        # an empty host map: nothing qualifies
        hm0 = golap.HostHashMap(key)
        hm0.build([(D["c_nation"], 99, 99)])
        e = ctx.star_run_hostdim(dims, cols["rev"], 1, hm0)
        assert (e[0].n_groups, e[0].rows_matched) == (0, 0)
        hm0.free()
    finally:
        ctx.free(dmask)
        hm.free()
        release(ctx, cols, maps, bufs)


def test_hostdim_refusals(golap, ctx):
    D = star_data(golap, 300_000, seed=3)
    cols, maps, bufs, dims, specs = setup(ctx, golap, D)
    hm = golap.HostHashMap(D["c_key"])
    try:
        with pytest.raises(golap.GolapError):  # host map not built
            ctx.star_run_hostdim(dims, cols["rev"], 1, hm)
        hm.build()
        with pytest.raises(golap.GolapError):  # no device dimension left
            ctx.star_run_hostdim(dims[1:2], cols["rev"], 0, hm)
        with pytest.raises(golap.GolapError):  # host dim out of range
            ctx.star_run_hostdim(dims, cols["rev"], 4, hm)
        # payload outside the host dim's declared group range
        hp = golap.HostHashMap(D["p_key"])
        hp.build(payload=np.full(D["p_key"].size, 5000, np.int32))
        with pytest.raises(golap.GolapError):
            ctx.star_run_hostdim(dims, cols["rev"], 3, hp)
        hp.free()
        r = ctx.star_run(dims, cols["rev"])  # the context still works
        assert r[0].n_groups >= 0
    finally:
        hm.free()
        release(ctx, cols, maps, bufs)
