"""Q1 streamed from column files (query1's SSD -> GPU pipeline, usecases/ssb/queries/01.hpp:
350-424, with skip_step for pruned blocks, 01.hpp:368-378): windows of chunks are read from
storage while earlier windows are scanned. The result equals the Q1 plan over the same columns
resident in HBM and the oracle; pruned chunks are never read."""
import numpy as np
import pytest

from gpu_util import q1_dataset, say
from test_gpu_q1 import date_map, release, upload_q1

pytestmark = pytest.mark.gpu


def save_all(hosts, tmp_path):
    paths = []
    for i, h in enumerate(hosts):
        p = tmp_path / f"q1_{i}.gcol"
        h.save(p)
        paths.append(p)
    return paths


@pytest.mark.parametrize("window", [1, 3, 0])
@pytest.mark.parametrize("codec,mm,page_rows", [(1, 16, 4096), (255, 4, 2048)])
def test_streamed_q1_equals_resident(orc, golap, ctx, tmp_path, window, codec, mm, page_rows):
    n = 1_500_007
    ds = q1_dataset(orc, n, seed=71)
    say("streamed q1 window", window, "codec", codec, "page_rows", page_rows)
    cols, hosts = upload_q1(ctx, golap, ds, codec=codec, page_rows=page_rows, chunk_rows=65536, lz4_min_match=mm)
    m, bufs = date_map(ctx, golap, ds, 1995)
    paths = save_all(hosts, tmp_path)
    try:
        want = ctx.q1_run(*cols, m, (4, 6), (30, 50))
        got, st = ctx.q1_run_files(paths, m, (4, 6), (30, 50), window_chunks=window)
        say("stream", st.as_dict())
        cpu, cnt = orc.q1_plan(ds["d_key"], [(ds["d_year"], 1995, 1995)], ds["date"], ds["disc"], ds["qty"],
                               ds["price"], (4, 6), (30, 50))
        assert (got.sum, got.matches) == (want.sum, want.matches) == (cpu, cnt)
        nch = hosts[0].arrays()["desc"].n_chunks
        assert (got.chunks_run, got.chunks_pruned, got.comp_bytes) == (nch, 0, want.comp_bytes)
        w = st.window_chunks
        assert w == (window or w) and st.windows == -(-nch // w) and st.windows_skipped == 0
        payload = sum(h.arrays()["desc"].payload_bytes for h in hosts)
        assert payload <= st.bytes_copied <= sum(h.arrays()["comp"].size for h in hosts)
        assert st.bytes_read >= st.bytes_copied and st.bytes_read % 4096 == 0
    finally:
        release(ctx, cols, m, bufs)


def test_streamed_q1_prunes_before_reading(orc, golap, ctx, tmp_path):
    """A price-sorted table: zone maps of the price column prune whole windows and runs inside
    windows; a host mask prunes more. Pruned chunks are not read and the result equals the
    unpruned resident plan."""
    ds = q1_dataset(orc, 2_000_000, seed=9)
    order = np.argsort(ds["price"], kind="stable")
    for k in ("date", "disc", "qty", "price"):
        ds[k] = np.ascontiguousarray(ds[k][order])
    cols, hosts = upload_q1(ctx, golap, ds, codec=1, page_rows=4096, chunk_rows=65536, lz4_min_match=16)
    m, bufs = date_map(ctx, golap, ds, 1994)
    paths = save_all(hosts, tmp_path)
    prng = (2_000_000, 5_000_000)
    try:
        full = ctx.q1_run(*cols, m, (1, 9), (1, 50), prng)
        pa = hosts[3].arrays()
        nch = pa["desc"].n_chunks
        zone = (pa["chunk_max"] >= prng[0]) & (pa["chunk_min"] <= prng[1])
        assert 0 < zone.sum() < nch
        got, st = ctx.q1_run_files(paths, m, (1, 9), (1, 50), prng, window_chunks=4)
        say("pruned stream", st.as_dict())
        assert (got.sum, got.matches) == (full.sum, full.matches)
        assert got.chunks_run == zone.sum() and got.chunks_pruned == nch - zone.sum()
        assert st.windows_skipped > 0 and st.windows + st.windows_skipped == -(-nch // 4)
        total = sum(h.arrays()["comp"].size for h in hosts)
        assert st.bytes_read < total
        # a host mask (every other chunk) on top: results equal the resident plan with that mask
        hm = np.zeros(nch, np.uint16)
        hm[::2] = 1
        d = ctx.to_device(hm)
        try:
            want = ctx.q1_run(*cols, m, (1, 9), (1, 50), prng, dev_chunk_mask=d)
        finally:
            ctx.free(d)
        got2, st2 = ctx.q1_run_files(paths, m, (1, 9), (1, 50), prng, host_chunk_mask=hm, window_chunks=5)
        assert (got2.sum, got2.matches) == (want.sum, want.matches)
        assert got2.chunks_run == int((zone & (hm > 0)).sum())
        assert st2.bytes_read < st.bytes_read
        # nothing kept: nothing read, an empty result
        got3, st3 = ctx.q1_run_files(paths, m, (1, 9), (1, 50), prng, host_chunk_mask=np.zeros(nch, np.uint16))
        assert (got3.sum, got3.matches, got3.chunks_run, st3.windows, st3.bytes_read) == (0, 0, 0, 0, 0)
        assert got3.chunks_pruned == nch
        # no pruning: every chunk read
        got4, st4 = ctx.q1_run_files(paths, m, (1, 9), (1, 50), prng, zone_prune=False, window_chunks=7)
        assert (got4.sum, got4.matches, got4.chunks_run) == (full.sum, full.matches, nch)
    finally:
        release(ctx, cols, m, bufs)


def test_streamed_q1_refusals(orc, golap, ctx, tmp_path):
    ds = q1_dataset(orc, 300_000, seed=5)
    cols, hosts = upload_q1(ctx, golap, ds, codec=1, page_rows=4096, chunk_rows=65536, lz4_min_match=16)
    m, bufs = date_map(ctx, golap, ds, 1995)
    paths = save_all(hosts, tmp_path)
    try:
        with pytest.raises(golap.GolapError):  # mask shorter than the chunk count
            ctx.q1_run_files(paths, m, (4, 6), (30, 50), host_chunk_mask=np.ones(2, np.uint16))
        short = golap.HostColumn.encode(ds["disc"][:100_000], codec=1, page_rows=4096, chunk_rows=65536)
        sp = tmp_path / "short.gcol"
        short.save(sp)
        with pytest.raises(golap.GolapError):  # not row-aligned
            ctx.q1_run_files([paths[0], sp, paths[2], paths[3]], m, (4, 6), (30, 50))
        with pytest.raises(golap.GolapError):
            ctx.q1_run_files([paths[0], tmp_path / "missing.gcol", paths[2], paths[3]], m, (4, 6), (30, 50))
        bad = tmp_path / "bad.gcol"
        bad.write_bytes(b"Y" + paths[1].read_bytes()[1:])
        with pytest.raises(golap.GolapError):
            ctx.q1_run_files([paths[0], bad, paths[2], paths[3]], m, (4, 6), (30, 50))
        # the context still works afterwards
        got, _ = ctx.q1_run_files(paths, m, (4, 6), (30, 50))
        want = ctx.q1_run(*cols, m, (4, 6), (30, 50))
        assert (got.sum, got.matches) == (want.sum, want.matches)
    finally:
        release(ctx, cols, m, bufs)


def test_streamed_q1_throughput(orc, golap, ctx, tmp_path):
    """40M rows (about 240 MB of LZ4 pages) streamed in default windows: reports the end-to-end
    rate (file reads, copies and scans overlapped) beside the kernels' share."""
    n = 40_000_000
    ds = q1_dataset(orc, n, seed=12)
    hosts = [golap.HostColumn.encode(ds[k], codec=1, page_rows=4096, chunk_rows=262144, lz4_min_match=16)
             for k in ("date", "disc", "qty", "price")]
    paths = save_all(hosts, tmp_path)
    m, bufs = date_map(ctx, golap, ds, 1995)
    try:
        got, st = ctx.q1_run_files(paths, m, (4, 6), (30, 50))
        cpu, cnt = orc.q1_plan(ds["d_key"], [(ds["d_year"], 1995, 1995)], ds["date"], ds["disc"], ds["qty"],
                               ds["price"], (4, 6), (30, 50), threads=4)
        assert (got.sum, got.matches) == (cpu, cnt)
        say("stream 40M", st.as_dict(), f"kernel_ms {got.kernel_ms:.2f}",
            f"{st.bytes_read / (st.total_ms * 1e-3) / 1e9:.2f} GB/s end to end",
            f"{n / (st.total_ms * 1e-3) / 1e9:.2f} G rows/s")
    finally:
        ctx.free_hashmap(m)
        for b in bufs:
            ctx.free(b)


@pytest.mark.parametrize("window", [1, 5, 0])
def test_streamed_star_equals_resident(orc, golap, ctx, tmp_path, window):
    """Star plans (query2/query3) streamed from column files in windows: the same groups as
    golap_star_run over the resident columns and the oracle; a host mask prunes before reading."""
    from test_gpu_star import build_map, star_data
    D = star_data(golap, 1_500_000, seed=41)
    enc = dict(page_rows=4096, chunk_rows=65536, lz4_min_match=4)
    names = ("date", "cust", "supp", "part", "rev")
    hosts = {n: golap.HostColumn.encode(D[n], codec=2 if n == "date" else 1, **enc) for n in names}
    cols = {n: ctx.upload(h) for n, h in hosts.items()}
    paths = {}
    for n, h in hosts.items():
        paths[n] = tmp_path / f"star_{n}.gcol"
        h.save(paths[n])
    maps, bufs = [], []
    for key, n_dim, preds, pay in ((D["d_key"], 2556, [], D["d_year"]),
                                   (D["c_key"], D["c_key"].size, [(D["c_nation"], 4, 9)], None),
                                   (D["s_key"], D["s_key"].size, [(D["s_region"], 1, 1)], None),
                                   (D["p_key"], D["p_key"].size, [(D["p_brand"], 1, 400)], D["p_brand"])):
        m, b = build_map(ctx, golap, key, n_dim, preds, pay)
        maps.append(m)
        bufs += b
    try:
        groups = [(1992, 7), (0, 0), (0, 0), (1, 1000)]
        dims = [(cols[n], maps[i], *groups[i]) for i, n in enumerate(names[:4])]
        want = ctx.star_run(dims, cols["rev"])
        fdims = [(paths[n], maps[i], *groups[i]) for i, n in enumerate(names[:4])]
        got = ctx.star_run_files(fdims, paths["rev"], window_chunks=window)
        say("streamed star window", window, got[4].as_dict())
        for a, b in zip(got[1:4], want[1:4]):
            assert np.array_equal(a, b)
        assert (got[0].rows_matched, got[0].total_sum, got[0].chunks_run, got[0].comp_bytes) == \
            (want[0].rows_matched, want[0].total_sum, want[0].chunks_run, want[0].comp_bytes)
        wk, ws, wc = orc.star_groupby(
            [(D["d_key"], [], D["d_year"]), (D["c_key"], [(D["c_nation"], 4, 9)], None),
             (D["s_key"], [(D["s_region"], 1, 1)], None), (D["p_key"], [(D["p_brand"], 1, 400)], D["p_brand"])],
            [D["date"], D["cust"], D["supp"], D["part"]], D["rev"], threads=4)
        assert np.array_equal(got[1], wk) and np.array_equal(got[2], ws) and np.array_equal(got[3], wc)
        # a host mask: the streamed plan equals the resident plan with the same device mask
        nch = hosts["rev"].arrays()["desc"].n_chunks
        hm = (np.arange(nch) % 3 == 0).astype(np.uint16)
        dm = ctx.to_device(hm)
        try:
            wantm = ctx.star_run(dims, cols["rev"], dev_chunk_mask=dm)
        finally:
            ctx.free(dm)
        gotm = ctx.star_run_files(fdims, paths["rev"], host_chunk_mask=hm, window_chunks=window)
        for a, b in zip(gotm[1:4], wantm[1:4]):
            assert np.array_equal(a, b)
        assert gotm[4].bytes_read < got[4].bytes_read
    finally:
        for c in cols.values():
            ctx.free_column(c)
        for m in maps:
            ctx.free_hashmap(m)
        for b in bufs:
            ctx.free(b)
