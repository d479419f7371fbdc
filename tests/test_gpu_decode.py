"""GPU page decoders (ChunkDecompressor / nvcomp decompress analogue, core/comp.cuh:171-195)
against the oracle's CPU decode of the same pages: bit-exact bytes, zero-padded last chunk."""
import numpy as np
import pytest

from conftest import rng
from gpu_util import load_liblz4_fixture, say

pytestmark = pytest.mark.gpu


def make(kind, n, dtype, seed=5):
    r = rng(seed)
    info = np.iinfo(dtype)
    if kind == "uniform":
        return r.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
    if kind == "lowcard":
        return r.integers(0, 11, n).astype(dtype)
    if kind == "sorted":
        return np.sort(r.integers(0, min(info.max, 3000), n)).astype(dtype)
    return np.repeat(r.integers(0, 7, n // 37 + 1), 37)[:n].astype(dtype)


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.int32, np.int64])
@pytest.mark.parametrize("kind", ["uniform", "lowcard", "sorted", "runs"])
@pytest.mark.parametrize("codec,mm", [(1, 4), (1, 16), (2, 4), (3, 4), (0, 4), (255, 4)])
def test_decode_matches_oracle(orc, golap, ctx, dtype, kind, codec, mm):
    n = 5 * 2048 + 333
    v = make(kind, n, dtype)
    say("decode", np.dtype(dtype).name, kind, "codec", codec, "min_match", mm, "rows", n)
    hc = golap.HostColumn.encode(v, codec=codec, page_rows=1024, chunk_rows=4096, lz4_min_match=mm)
    want = orc.decode_column(hc.arrays())
    col = ctx.upload(hc)
    try:
        got, ms = ctx.decompress(col)
        assert np.array_equal(got, want)
        d = ctx.column_desc(col)
        got2, _ = ctx.decompress(col, 1, d.n_chunks)  # chunk sub-range
        assert np.array_equal(got2, want[d.chunk_rows * d.elem:])
    finally:
        ctx.free_column(col)


def test_decode_liblz4_made_pages(orc, golap, ctx):
    data, page_rows, pages = load_liblz4_fixture()
    say("decode liblz4 fixture pages", len(pages))
    hc = golap.HostColumn.assemble(golap.U8, data.size, page_rows, page_rows * 4, pages, [1] * len(pages))
    want = orc.decode_column(hc.arrays())
    assert want[:data.size].tobytes() == data.tobytes()
    col = ctx.upload(hc)
    try:
        got, _ = ctx.decompress(col)
        assert np.array_equal(got, want)
    finally:
        ctx.free_column(col)


def test_large_column_decode(orc, golap, ctx):
    v = make("lowcard", 3_000_000, np.int32, seed=9)
    say("decode large lowcard int32 column", v.size)
    hc = golap.HostColumn.encode(v, codec=1, page_rows=4096, chunk_rows=262144, lz4_min_match=4)
    col = ctx.upload(hc)
    try:
        got, ms = ctx.decompress(col)
        assert got[:v.nbytes].tobytes() == v.tobytes()
        assert not got[v.nbytes:].any()
    finally:
        ctx.free_column(col)


@pytest.mark.parametrize("damage", ["truncate", "bad_offset"])
def test_corrupt_page_reports_ecorrupt(orc, golap, ctx, damage):
    data = np.frombuffer((b"abcdefgh" * 1024), np.uint8)
    from conftest import Lz4Lib
    try:
        blk = Lz4Lib().compress(data.tobytes())
    except OSError:
        pytest.skip("liblz4 missing")
    if damage == "truncate":
        blk = blk[:len(blk) // 2]
    else:
        b = bytearray(blk)
        b[b.index(ord("h")) + 1:b.index(ord("h")) + 3] = b"\xff\xff"  # offset beyond the output
        blk = bytes(b)
    assert orc.lz4_decode(blk, data.size) is None
    say("corrupt page", damage)
    hc = golap.HostColumn.assemble(golap.U8, data.size, data.size, data.size, [blk], [1])
    col = ctx.upload(hc)
    try:
        with pytest.raises(golap.GolapError) as e:
            ctx.decompress(col)
        assert e.value.status == golap.ECORRUPT
        # the context stays usable after the error
        ok = golap.HostColumn.encode(data, codec=1, page_rows=8192, chunk_rows=8192)
        c2 = ctx.upload(ok)
        got, _ = ctx.decompress(c2)
        assert got.tobytes() == data.tobytes()
        ctx.free_column(c2)
    finally:
        ctx.free_column(col)
