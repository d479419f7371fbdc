"""Caller streams and events (CStream / CEvent and the Op protocol's step(stream, parent_event),
core/core.hpp:94-124, 192-307), the chunk decode enqueued on a stream
(ChunkDecompressor::step on a worker's stream, core/comp.cuh:171-195) and the device-resident
group-by accumulator (HashAggregate, core/dev_structs.cuh:157-253), against the oracle."""
import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu


def test_async_decode_on_two_streams_matches_oracle(orc, golap, ctx):
    r = np.random.default_rng(81)
    n = 700_001
    a = (19920101 + r.integers(0, 2556, n)).astype(np.int32)     # LZ4I pages with matches
    b = np.repeat(r.integers(0, 50, n // 7 + 1), 7)[:n].astype(np.int32)
    say("async decode two streams rows", n)
    cols, bufs, streams = [], [], []
    try:
        for v in (a, b):
            hc = golap.HostColumn.encode(v, codec=1, page_rows=4096, chunk_rows=65536, lz4_min_match=4)
            cols.append((ctx.upload(hc), orc.decode_column(hc.arrays())))
        for c, want in cols:
            bufs.append(ctx.alloc(want.size))
            streams.append(ctx.stream())
        ev = ctx.event()
        ctx.event_record(ev, None)  # the parent: the context's stream
        for (c, _), buf, s in zip(cols, bufs, streams):
            ctx.stream_wait_event(s, ev)
            ctx.decompress_async(c, buf, stream=s)
        for s in streams:
            ctx.stream_sync(s)
        for (c, want), buf in zip(cols, bufs):
            assert np.array_equal(ctx.to_host(buf, want.size, np.uint8), want)
        # the context's stream waits for a worker's event before reading its buffer
        e2 = ctx.event()
        ctx.decompress_async(cols[0][0], bufs[0], 1, 3, stream=streams[0])
        ctx.event_record(e2, streams[0])
        ctx.stream_wait_event(None, e2)
        got = ctx.to_host(bufs[0], 2 * 65536 * 4, np.uint8)
        assert np.array_equal(got, cols[0][1][65536 * 4:3 * 65536 * 4])
        ctx.free_event(e2)
        ctx.free_event(ev)
    finally:
        for s in streams:
            ctx.free_stream(s)
        for buf in bufs:
            ctx.free(buf)
        for c, _ in cols:
            ctx.free_column(c)


def test_async_decode_reports_corrupt_page_at_stream_sync(orc, golap, ctx):
    data = np.frombuffer(b"abcdefgh" * 1024, np.uint8)
    a = golap.HostColumn.encode(np.tile(data, 4), codec=1, page_rows=8192, chunk_rows=8192).arrays()
    pg = bytearray(a["comp"][int(a["page_off"][0]):int(a["page_off"][0]) + int(a["page_csize"][0])].tobytes())
    pg = pg[:len(pg) // 2]  # truncated page
    assert orc.page_decode(int(a["page_codec"][0]), bytes(pg), 8192)[0] != 8192
    hc = golap.HostColumn.assemble(golap.U8, 8192, 8192, 8192, [bytes(pg)], [int(a["page_codec"][0])])
    col = ctx.upload(hc)
    buf = ctx.alloc(8192)
    s = ctx.stream()
    try:
        ctx.decompress_async(col, buf, stream=s)  # enqueues; the error shows at the sync
        with pytest.raises(golap.GolapError) as e:
            ctx.stream_sync(s)
        assert e.value.status == golap.ECORRUPT
        ctx.stream_sync(s)  # the stream's error word was cleared
        ok = ctx.upload(golap.HostColumn.encode(data, codec=1, page_rows=8192, chunk_rows=8192))
        ctx.decompress_async(ok, buf, stream=s)
        ctx.stream_sync(s)
        assert ctx.to_host(buf, 8192, np.uint8).tobytes() == data.tobytes()
        ctx.free_column(ok)
    finally:
        ctx.free_stream(s)
        ctx.free(buf)
        ctx.free_column(col)


def test_hashagg_accumulates_across_calls_and_streams(orc, golap, ctx):
    r = np.random.default_rng(82)
    n = 1_000_000
    k0 = r.integers(-3, 8, n).astype(np.int32)
    k1 = r.integers(0, 300, n).astype(np.int16)
    v = r.integers(-1000, 100000, n).astype(np.int32)
    say("hashagg rows", n)
    d0, d1, dv = ctx.to_device(k0), ctx.to_device(k1), ctx.to_device(v)
    streams = [ctx.stream(), ctx.stream()]
    h = ctx.hashagg([(golap.I32, -3, 11), (golap.I16, 0, 300)])
    try:
        half = n // 2  # the two halves on two streams, the second half in two calls
        ctx.hashagg_add(h, [d0, d1], dv, golap.I32, half, stream=streams[0])
        q = half + (n - half) // 2
        ctx.hashagg_add(h, [d0 + half * 4, d1 + half * 2], dv + half * 4, golap.I32, q - half, stream=streams[1])
        ctx.hashagg_add(h, [d0 + q * 4, d1 + q * 2], dv + q * 4, golap.I32, n - q, stream=streams[1])
        for s in streams:
            ctx.stream_sync(s)
        keys, sums, cnt = ctx.hashagg_results(h, 2, 11 * 300)
        wk, ws, wc = orc.groupby([k0, k1.astype(np.int32)], v.astype(np.int64))
        assert np.array_equal(keys, wk) and np.array_equal(sums, ws) and np.array_equal(cnt, wc)
        # a key out of its range is reported by the stream it ran on
        bad = ctx.to_device(np.full(16, 9, np.int32))
        ctx.hashagg_add(h, [bad, d1], dv, golap.I32, 16, stream=streams[0])
        with pytest.raises(golap.GolapError):
            ctx.stream_sync(streams[0])
        ctx.free(bad)
    finally:
        ctx.free_hashagg(h)
        for s in streams:
            ctx.free_stream(s)
        for d in (d0, d1, dv):
            ctx.free(d)
