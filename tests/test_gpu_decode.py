"""GPU page decoders (ChunkDecompressor / nvcomp decompress analogue, core/comp.cuh:171-195)
against the oracle's CPU decode of the same pages: bit-exact bytes, zero-padded last chunk."""
import numpy as np
import pytest

from conftest import LZ4I, lz4_block, rng
from test_lz4i_cpu import make_page
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


@pytest.mark.parametrize("dtype,page_rows", [(np.uint8, 4096), (np.uint8, 16384), (np.int32, 4096)])
def test_decode_long_match_chains(orc, golap, ctx, dtype, page_rows):
    """Pages whose matches copy from other matches in chains tens of links deep (small Zipf
    values: the trip table's passenger count), where most matches are left pending by the ordered
    pass and resolved by pointer jumping; the typed case resolves whole words."""
    r = rng(17)
    n = 40 * page_rows + 123
    v = np.minimum(r.zipf(1.2, n), 6).astype(dtype)
    say("decode zipf chains", np.dtype(dtype).name, "page_rows", page_rows, "rows", n)
    for unit in (1, 0):
        hc = golap.HostColumn.encode(v, codec=1, page_rows=page_rows, chunk_rows=4 * page_rows, lz4_min_match=4,
                                     lz4_unit=unit)
        want = orc.decode_column(hc.arrays())
        assert want[:v.nbytes].tobytes() == v.tobytes()
        col = ctx.upload(hc)
        try:
            got, _ = ctx.decompress(col)
            assert np.array_equal(got, want)
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


def _lz4i_cases():
    r = rng(61)
    dates = (19920101 + r.integers(0, 2556, 4096 * 12 + 5)).astype(np.int32)
    yield "dates_4096", dates, 4096
    yield "dates_8192", dates, 8192
    yield "lowcard_u8_4096", r.integers(0, 11, 4096 * 12).astype(np.uint8), 4096
    # long literal runs and long overlapping matches
    runs = np.repeat(r.integers(0, 3, 3000), r.integers(1, 200, 3000))[:4096 * 8].astype(np.int32)
    yield "runs_i32_4096", runs, 4096
    # few matches over a nearly incompressible page: the groups do not fit in place (serial)
    planted = r.integers(0, 2**31, 4096 * 8).astype(np.int32)
    for k in range(0, planted.size - 64, 200):  # ten or more sequences per page
        planted[k + 40:k + 48] = planted[k:k + 8]
    yield "planted_i32_4096", planted, 4096


@pytest.mark.parametrize("name,values,page_rows", list(_lz4i_cases()), ids=[c[0] for c in _lz4i_cases()])
def test_decode_lz4i_pages_match_oracle(orc, golap, ctx, name, values, page_rows):
    """LZ4I pages through the decompress kernel's parallel parse (and its serial fallback),
    against the oracle: as the encoder made them, and re-indexed with 16- and 64-sequence
    segments (more than 64 segments per page: lanes walk several)."""
    say("decode lz4i", name, "page_rows", page_rows)
    hc = golap.HostColumn.encode(values, codec=1, page_rows=page_rows, chunk_rows=page_rows * 2, lz4_min_match=4)
    a = hc.arrays()
    d = a["desc"]
    assert (a["page_codec"] == LZ4I).sum() >= d.n_pages - 1
    variants = [hc]
    for sl in (4, 6):
        pages = []
        for p in range(d.n_pages):
            off, cs, code = int(a["page_off"][p]), int(a["page_csize"][p]), int(a["page_codec"][p])
            pg = a["comp"][off:off + cs].tobytes()
            pages.append(make_page(lz4_block(pg, code), sl) if code == LZ4I else pg)
        codes = [LZ4I if c == LZ4I else int(c) & 0x0F for c in a["page_codec"]]
        variants.append(golap.HostColumn.assemble(golap.TYPE_OF[values.dtype], values.size, page_rows,
                                                  page_rows * 2, pages, codes))
    for v in variants:
        want = orc.decode_column(v.arrays())
        assert want is not None and want[:values.nbytes].tobytes() == values.tobytes()
        col = ctx.upload(v)
        try:
            got, _ = ctx.decompress(col)
            assert np.array_equal(got, want)
        finally:
            ctx.free_column(col)


@pytest.mark.parametrize("damage", ["entry", "count", "pad"])
@pytest.mark.parametrize("rows", [4096, 64])
def test_corrupt_lz4i_index_reports_ecorrupt(orc, golap, ctx, damage, rows):
    """A damaged index is refused by the GPU decoders as by the oracle: by the parallel parse
    (4,096-row page) and by the serial decoder (a 64-row page of few sequences)."""
    r = rng(62)
    v = (19920101 + r.integers(0, 2556, rows)).astype(np.int32)
    a = golap.HostColumn.encode(v, codec=1, page_rows=rows, chunk_rows=rows, lz4_min_match=4).arrays()
    assert a["page_codec"][0] == LZ4I
    b = bytearray(a["comp"][int(a["page_off"][0]):int(a["page_off"][0]) + int(a["page_csize"][0])].tobytes())
    nseq = b[0] | (b[1] << 8)
    if damage == "entry":
        if nseq <= 16:
            b[2] = 1  # two-sequence segments: the header grows, the block no longer starts there
        else:
            b[4] ^= 1
    elif damage == "count":
        b[0], b[1] = (nseq - 1) & 0xFF, (nseq - 1) >> 8
    else:
        b[3] = 9
    assert orc.page_decode(LZ4I, bytes(b), v.nbytes)[0] == -1
    say("corrupt lz4i", damage, "rows", rows, "sequences", nseq)
    hc = golap.HostColumn.assemble(golap.I32, rows, rows, rows, [bytes(b)], [LZ4I])
    col = ctx.upload(hc)
    try:
        with pytest.raises(golap.GolapError) as e:
            ctx.decompress(col)
        assert e.value.status == golap.ECORRUPT
    finally:
        ctx.free_column(col)


@pytest.mark.parametrize("dtype", [np.int16, np.int32, np.int64])
@pytest.mark.parametrize("page_rows", [1024, 4096])
def test_decode_typed_pages_match_values(orc, golap, ctx, dtype, page_rows):
    """Typed LZ4 pages (sequences in whole units of the element width, 4 bytes for int64, as the
    reference's typed nvcomp LZ4 codes them): the cooperative decoder copies whole words; the
    result equals the values, the oracle and the byte-granular pages of the same values."""
    r = rng(71)
    n = page_rows * 5 + 77
    v = (np.repeat(r.integers(0, 1000, n), r.integers(1, 6, n))[:n] * 7 + 12345).astype(dtype)  # runs: many matches
    say("decode typed", np.dtype(dtype).name, "page_rows", page_rows)
    for unit in (0, 1):
        hc = golap.HostColumn.encode(v, codec=1, page_rows=page_rows, chunk_rows=page_rows * 2, lz4_min_match=4,
                                     lz4_unit=unit)
        a = hc.arrays()
        lz4i = [p for p in range(a["desc"].n_pages) if a["page_codec"][p] == LZ4I]
        assert len(lz4i) >= 5
        first = a["comp"][int(a["page_off"][lz4i[0]]) + 3]
        assert first == ((min(v.itemsize, 4).bit_length() - 1) if unit == 0 else 0)
        want = orc.decode_column(a)
        assert want[:v.nbytes].tobytes() == v.tobytes()
        col = ctx.upload(hc)
        try:
            got, _ = ctx.decompress(col)
            assert np.array_equal(got, want)
        finally:
            ctx.free_column(col)


def test_false_unit_claim_reports_ecorrupt(orc, golap, ctx):
    """A byte-granular page that claims 4-byte units is refused by the GPU decoders as by the
    oracle (the cooperative decoder checks every sequence's lengths and offset)."""
    r = rng(72)
    v = (19920101 + r.integers(0, 2556, 4096)).astype(np.int32)
    a = golap.HostColumn.encode(v, codec=1, page_rows=4096, chunk_rows=4096, lz4_min_match=4, lz4_unit=1).arrays()
    b = bytearray(a["comp"][int(a["page_off"][0]):int(a["page_off"][0]) + int(a["page_csize"][0])].tobytes())
    assert b[3] == 0 and orc.page_decode(LZ4I, bytes(b), v.nbytes)[0] == v.nbytes
    b[3] = 2
    assert orc.page_decode(LZ4I, bytes(b), v.nbytes)[0] == -1
    hc = golap.HostColumn.assemble(golap.I32, 4096, 4096, 4096, [bytes(b)], [LZ4I])
    col = ctx.upload(hc)
    try:
        with pytest.raises(golap.GolapError) as e:
            ctx.decompress(col)
        assert e.value.status == golap.ECORRUPT
    finally:
        ctx.free_column(col)
