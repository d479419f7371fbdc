"""Join-key pruning (usecases/ssb/queries/01.hpp:246-270): collect_hashmap_keys
(core/dev_structs.cuh:136-150) gathers the keys a built date map holds, and
check_mmmeta_multiple_pts (core/metadata.cuh:550-559) keeps a chunk of the fact's date column
iff one of them lies in the chunk's [min, max]. Checked against a brute-force restatement of
those two loops over the host zone maps, and by the invariant pruned result == unpruned."""
import numpy as np
import pytest

from gpu_util import q1_dataset, say
from test_gpu_q1 import date_map, release, upload_q1

pytestmark = pytest.mark.gpu


def brute_points(mins, maxs, keys):
    """check_mmmeta_multiple_pts as the reference loops it: every chunk against every key."""
    out = np.zeros(mins.size, np.uint16)
    for k in keys:
        out |= ((k >= mins) & (k <= maxs)).astype(np.uint16)
    return out


@pytest.fixture(scope="module")
def sorted_q1(orc, golap, ctx):
    ds = q1_dataset(orc, 2_000_000, seed=31)
    order = np.argsort(ds["date"], kind="stable")
    for k in ("date", "disc", "qty", "price"):
        ds[k] = np.ascontiguousarray(ds[k][order])
    cols, hosts = upload_q1(ctx, golap, ds, codec=1, page_rows=4096, chunk_rows=65536, lz4_min_match=16)
    yield ds, cols, hosts
    for c in cols:
        ctx.free_column(c)


@pytest.mark.parametrize("year", [1992, 1995, 1998, 1990])
def test_hashmap_keys_are_the_qualifying_build_keys(golap, ctx, sorted_q1, year):
    ds, cols, hosts = sorted_q1
    m, bufs = date_map(ctx, golap, ds, year)
    keys = ctx.alloc(4 * ds["d_key"].size)
    try:
        n = ctx.hashmap_keys(m, keys, ds["d_key"].size)
        got = ctx.to_host(keys, n, np.int32) if n else np.zeros(0, np.int32)
        want = np.sort(ds["d_key"][ds["d_year"] == year]).astype(np.int32)
        say("year", year, "keys", n)
        assert np.array_equal(got, want)
        if n > 1:  # a short output buffer is an error that still reports the count
            with pytest.raises(golap.GolapError):
                ctx.hashmap_keys(m, keys, n - 1)
    finally:
        ctx.free(keys)
        release(ctx, [], m, bufs)


def test_points_check_matches_brute_force(golap, ctx, sorted_q1):
    ds, cols, hosts = sorted_q1
    a = hosts[0].arrays()
    mins, maxs = a["chunk_min"], a["chunk_max"]
    nch = mins.size
    rng = np.random.default_rng(5)
    flags = ctx.alloc(nch * 2)
    keys = ctx.alloc(4 * 4096)
    try:
        cases = [np.zeros(0, np.int32),                                     # no key: every chunk pruned
                 np.array([mins[3], maxs[7]], np.int32),                      # inclusive bounds
                 np.array([mins[0] - 1, maxs[-1] + 1], np.int32),             # just outside the table
                 np.sort(rng.integers(int(mins.min()) - 5, int(maxs.max()) + 5, 300)).astype(np.int32),
                 np.sort(ds["d_key"]).astype(np.int32)]
        for ks in cases:
            if ks.size:
                ctx.h2d(keys, ks)
            ctx.minmax_check_points(cols[0], keys, ks.size, flags)
            got = ctx.to_host(flags, nch, np.uint16)
            assert np.array_equal(got, brute_points(mins, maxs, ks)), ks[:4]
    finally:
        ctx.free(flags)
        ctx.free(keys)


def test_join_pruned_q1_equals_unpruned(orc, golap, ctx, sorted_q1):
    """01.hpp:246-270 end to end: the date map's keys prune the date-sorted fact table; the
    pruned plan equals the unpruned one and the CPU plan."""
    ds, cols, hosts = sorted_q1
    m, bufs = date_map(ctx, golap, ds, 1995)
    nch = ctx.column_desc(cols[0]).n_chunks
    flags = ctx.alloc(nch * 2)
    f2 = ctx.alloc(nch * 2)
    keys = ctx.alloc(4 * ds["d_key"].size)
    try:
        n = ctx.hashmap_keys(m, keys, ds["d_key"].size)
        ctx.minmax_check_points(cols[0], keys, n, flags)
        ctx.minmax_check(cols[1], 4, 6, f2)
        ctx.combine_and(flags, f2, nch)
        kept = int(ctx.to_host(flags, nch, np.uint16).sum())
        assert 0 < kept < nch
        pruned = ctx.q1_run(*cols, m, (4, 6), (30, 50), dev_chunk_mask=flags)
        full = ctx.q1_run(*cols, m, (4, 6), (30, 50))
        assert (pruned.sum, pruned.matches) == (full.sum, full.matches)
        assert pruned.chunks_pruned == nch - kept
        cpu, cnt = orc.q1_plan(ds["d_key"], [(ds["d_year"], 1995, 1995)], ds["date"], ds["disc"], ds["qty"],
                               ds["price"], (4, 6), (30, 50))
        assert (full.sum, full.matches) == (cpu, cnt)
    finally:
        ctx.free(keys)
        release(ctx, [], m, bufs + [flags, f2])
