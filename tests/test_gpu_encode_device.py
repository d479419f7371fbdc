"""prepare_compressed's compression on the device (ChunkCompressor in the chain of
core/apps.cuh:148-255): values already in HBM coded by golap_encode_device give a device column
byte-identical to the host encoder's (golap_encode + upload): page bytes, catalog, layout and
zone maps, for every codec, type, page size and LZ4 minimum match, ragged and tiny inputs."""
import time

import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu

CODECS = {"raw": 0, "lz4": 1, "dbp": 2, "rle": 3, "for": 4, "auto": 255}


def pattern(kind, n, dtype, seed):
    rng = np.random.default_rng(seed)
    info = np.iinfo(dtype)
    if kind == "random":  # full range: LZ4 literal runs, FOR/DBP too wide for 64-bit types
        return rng.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
    if kind == "sorted":  # DBP
        lo = max(info.min, -1000)
        return np.clip(lo + np.cumsum(rng.integers(0, 3, n)), info.min, info.max).astype(dtype)
    if kind == "runs":  # RLE
        return np.repeat(rng.integers(max(info.min, -50), min(info.max, 50), n // 37 + 1), 37)[:n].astype(dtype)
    if kind == "lowrange":  # FOR
        return rng.integers(max(info.min, 100), min(info.max, 160), n).astype(dtype)
    if kind == "text":  # LZ4 matches, overlapping ones included
        motif = rng.integers(0, min(info.max, 200), 29).astype(dtype)
        v = np.tile(motif, n // 29 + 1)[:n].copy()
        noise = rng.random(n) < 0.02
        v[noise] = rng.integers(0, min(info.max, 200), int(noise.sum())).astype(dtype)
        return v
    if kind == "const":
        return np.full(n, min(info.max, 7), dtype)
    if kind == "extremes":  # deltas that leave int64 (DBP refused), FOR refused
        return rng.choice(np.array([info.min, info.max, 0], dtype=dtype), n)
    raise ValueError(kind)


def compare(ctx, golap, values, type_, **kw):
    n = values.size
    host = golap.HostColumn.encode(values, **kw)
    want = host.arrays()
    dv = ctx.to_device(values) if n else None
    col = None
    try:
        col, ms = ctx.encode_device(dv, n, type_, **kw)
        d = ctx.column_desc(col)
        assert (d.n_rows, d.n_pages, d.n_chunks) == (n, want["desc"].n_pages, want["desc"].n_chunks)
        assert d.comp_bytes == want["comp"].size and d.payload_bytes == want["desc"].payload_bytes
        for key, dt, cnt in (("page_off", np.uint64, d.n_pages), ("page_csize", np.uint32, d.n_pages),
                             ("page_codec", np.uint8, d.n_pages), ("chunk_min", np.int64, d.n_chunks),
                             ("chunk_max", np.int64, d.n_chunks)):
            got = ctx.to_host(getattr(d, key), cnt, dt) if cnt else np.zeros(0, dt)
            assert np.array_equal(got, want[key]), key
        comp = ctx.to_host(d.comp, d.comp_bytes, np.uint8)
        if not np.array_equal(comp, want["comp"]):
            bad = np.flatnonzero(comp != want["comp"])
            page = np.searchsorted(want["page_off"].astype(np.int64), bad[0], side="right") - 1
            raise AssertionError(f"comp differs at byte {bad[0]} ({bad.size} bytes), page {page} "
                                 f"codec {want['page_codec'][page]:#x}")
        return want["page_codec"], ms
    finally:
        if col:
            ctx.free_column(col)
        if dv:
            ctx.free(dv)


@pytest.mark.parametrize("codec", list(CODECS))
@pytest.mark.parametrize("dtype", [np.int32, np.uint8, np.int64, np.uint16])
def test_encode_device_matches_host(golap, ctx, codec, dtype):
    type_ = golap.TYPE_OF[np.dtype(dtype)]
    for i, kind in enumerate(["random", "sorted", "runs", "lowrange", "text", "const", "extremes"]):
        v = pattern(kind, 150_001, dtype, seed=i)
        say("encode_device", codec, np.dtype(dtype).name, kind)
        codes, _ = compare(ctx, golap, v, type_, codec=CODECS[codec], page_rows=4096, chunk_rows=65536,
                           lz4_min_match=4)
        if codec == "auto":
            assert len(set(codes & 0x0F)) >= 1


@pytest.mark.parametrize("page_rows,chunk_rows,min_match", [(64, 128, 4), (1024, 8192, 16), (2048, 2048, 6)])
def test_encode_device_page_shapes(golap, ctx, page_rows, chunk_rows, min_match):
    for codec in CODECS.values():
        for kind in ("text", "random", "sorted"):
            v = pattern(kind, 70_000 + page_rows // 3, np.int32, seed=page_rows)
            compare(ctx, golap, v, golap.I32, codec=codec, page_rows=page_rows, chunk_rows=chunk_rows,
                    lz4_min_match=min_match)


def test_encode_device_tiny_and_empty(golap, ctx):
    for n in (0, 1, 5, 12, 13, 64, 65):
        for codec in CODECS.values():
            v = pattern("text", n, np.uint8, seed=n)
            compare(ctx, golap, v, golap.U8, codec=codec, page_rows=64, chunk_rows=128, lz4_min_match=4)


def test_encode_device_large_raw_rle_pages(golap, ctx):
    v = pattern("runs", 300_000, np.int32, seed=3)
    for codec in (0, 3):
        compare(ctx, golap, v, golap.I32, codec=codec, page_rows=65536, chunk_rows=131072, lz4_min_match=4)


def test_encode_device_refusals(golap, ctx):
    v = ctx.to_device(np.arange(1000, dtype=np.int32))
    try:
        with pytest.raises(golap.GolapError):  # LZ4 pages above 64 KiB
            ctx.encode_device(v, 1000, golap.I32, codec=1, page_rows=32768, chunk_rows=32768)
        with pytest.raises(golap.GolapError):  # unaligned values
            ctx.encode_device(v + 4, 990, golap.I32, codec=0)
        with pytest.raises(golap.GolapError):
            ctx.encode_device(v, 1000, golap.I32, codec=9)
        with pytest.raises(golap.GolapError):
            ctx.encode_device(v, 1000, golap.I32, page_rows=100)
    finally:
        ctx.free(v)


def test_encode_device_decodes_and_runs_q1(orc, golap, ctx):
    """Columns coded on the device serve the Q1 plan like host-coded ones, and their rate."""
    from gpu_util import q1_dataset
    from test_gpu_q1 import date_map, release
    n = 20_000_000
    ds = q1_dataset(orc, n, seed=21)
    cols, rates = [], []
    for name in ("date", "disc", "qty", "price"):
        v = ds[name]
        dv = ctx.to_device(v)
        t = time.perf_counter()
        h = golap.HostColumn.encode(v, codec=1, page_rows=4096, chunk_rows=262144, lz4_min_match=16)
        host_s = time.perf_counter() - t
        col, ms = ctx.encode_device(dv, n, golap.TYPE_OF[v.dtype], codec=1, page_rows=4096, chunk_rows=262144,
                                    lz4_min_match=16)
        ctx.free(dv)
        cols.append(col)
        rates.append((name, v.nbytes / (ms * 1e-3) / 1e9, v.nbytes / host_s / 1e9))
        d = ctx.column_desc(col)
        assert np.array_equal(ctx.to_host(d.comp, d.comp_bytes, np.uint8), h.arrays()["comp"])
    say("encode_device LZ4 GB/s (device, host):", [(a, round(b, 1), round(c, 2)) for a, b, c in rates])
    m, bufs = date_map(ctx, golap, ds, 1995)
    try:
        r = ctx.q1_run(*cols, m, (4, 6), (30, 50))
        want, cnt = orc.q1_plan(ds["d_key"], [(ds["d_year"], 1995, 1995)], ds["date"], ds["disc"], ds["qty"],
                                ds["price"], (4, 6), (30, 50), threads=4)
        assert (r.sum, r.matches) == (want, cnt)
    finally:
        release(ctx, cols, m, bufs)


@pytest.mark.parametrize("dtype", [np.int16, np.int32, np.int64])
def test_encode_device_typed_matches_host(golap, ctx, dtype):
    """Typed LZ4 (sequences in whole units of the element width): the device encoder's pages,
    catalog and zone maps equal the host encoder's byte for byte."""
    type_ = golap.TYPE_OF[np.dtype(dtype)]
    for i, kind in enumerate(["runs", "lowrange", "text", "random"]):
        v = pattern(kind, 90_001, dtype, seed=40 + i)
        say("encode_device typed", np.dtype(dtype).name, kind)
        for codec in ("lz4", "auto"):
            compare(ctx, golap, v, type_, codec=CODECS[codec], page_rows=4096, chunk_rows=65536, lz4_min_match=4,
                    lz4_unit=0)
