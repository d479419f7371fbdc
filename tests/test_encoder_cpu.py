"""Host write path (prepare_compressed analogue, core/apps.cuh:148-255): the product encoder's
pages decoded by the oracle and by liblz4, catalog invariants of core/test/apps_tests.cu:12-65,
and zone maps against the oracle's MinMaxMeta restatement."""
import numpy as np
import pytest

from conftest import lz4_block, rng

PAGE_LITERAL_ONLY = 0x80


def gen(kind, n, dtype, seed=3):
    r = rng(seed)
    info = np.iinfo(dtype)
    if kind == "uniform":
        return r.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
    if kind == "lowcard":
        return r.integers(0, 11, n).astype(dtype)
    if kind == "sorted":
        return np.sort(r.integers(0, max(2, min(info.max, 5000)), n)).astype(dtype)
    if kind == "runs":
        return np.repeat(r.integers(0, 7, n // 50 + 1), 50)[:n].astype(dtype)
    if kind == "const":
        return np.full(n, 42, dtype=dtype)
    raise ValueError(kind)


def check_roundtrip(orc, golap, values, **kw):
    hc = golap.HostColumn.encode(values, **kw)
    a = hc.arrays()
    d = a["desc"]
    out = orc.decode_column(a)
    assert out is not None
    raw = values.tobytes()
    assert out[:len(raw)].tobytes() == raw
    assert not out[len(raw):].any()  # zero-padded tail of the last chunk
    return hc, a


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.int32, np.uint32, np.int64])
@pytest.mark.parametrize("kind", ["uniform", "lowcard", "sorted", "runs", "const"])
@pytest.mark.parametrize("codec", [1, 2, 3, 4, 255])
def test_encode_decode_roundtrip(orc, golap, dtype, kind, codec):
    n = 3 * 4096 + 777  # ragged last page and chunk
    values = gen(kind, n, dtype)
    check_roundtrip(orc, golap, values, codec=codec, page_rows=1024, chunk_rows=4096, threads=3)


@pytest.mark.parametrize("min_match", [4, 8, 16, 64])
def test_lz4_pages_decode_with_liblz4(orc, golap, lz4lib, min_match):
    values = gen("lowcard", 20000, np.uint8)
    hc, a = check_roundtrip(orc, golap, values, codec=1, page_rows=2048, chunk_rows=8192,
                            lz4_min_match=min_match)
    d = a["desc"]
    raw = values.tobytes()
    for p in range(d.n_pages):
        off, cs = int(a["page_off"][p]), int(a["page_csize"][p])
        rows = min(d.page_rows, d.n_rows - p * d.page_rows)
        page = lz4_block(a["comp"][off:off + cs], a["page_codec"][p])
        assert lz4lib.decompress_partial(page, rows) == raw[p * 2048:p * 2048 + rows]


def test_literal_only_pages_are_aligned_in_place(golap):
    values = gen("uniform", 50000, np.int32)
    a = golap.HostColumn.encode(values, codec=1, page_rows=4096, chunk_rows=16384).arrays()
    d = a["desc"]
    raw = values.view(np.uint8)
    for p in range(d.n_pages):
        assert a["page_codec"][p] == 1 | PAGE_LITERAL_ONLY  # iid 32-bit values: no matches
        rows = min(4096, d.n_rows - p * 4096)
        nb = rows * 4
        hdr = 1 if nb < 15 else 2 + (nb - 15) // 255
        start = int(a["page_off"][p]) + hdr
        assert start % 16 == 0
        assert int(a["page_csize"][p]) == hdr + nb
        assert np.array_equal(a["comp"][start:start + nb], raw[p * 16384:p * 16384 + nb])


def test_catalog_invariants(golap):
    # apps_tests.cu:12-65: block count == ceil(rows*size/CHUNK), tuples per block, contiguous
    # non-overlapping blocks
    n, page, chunk = 1_000_003, 1024, 65536
    values = gen("uniform", n, np.int32)
    a = golap.HostColumn.encode(values, codec=1, page_rows=page, chunk_rows=chunk).arrays()
    d = a["desc"]
    assert d.n_chunks == -(-n * 4 // (chunk * 4))
    assert d.n_pages == -(-n // page)
    ends = a["page_off"] + a["page_csize"]
    assert np.all(a["page_off"][1:] >= ends[:-1])
    assert np.all(a["page_off"][1:] - ends[:-1] < 16 + 16)  # only alignment gaps
    assert d.payload_bytes == int(a["page_csize"].sum())
    assert d.uncomp_bytes == n * 4


def test_zone_maps_match_oracle(orc, golap):
    values = gen("uniform", 300_000, np.int32)
    a = golap.HostColumn.encode(values, codec=1, page_rows=1024, chunk_rows=65536).arrays()
    mins, maxs = orc.minmax_build(values, 65536)
    assert np.array_equal(a["chunk_min"], mins)
    assert np.array_equal(a["chunk_max"], maxs)


def test_closed_forms(golap):
    # constant page -> one RLE run: header 16 + value padded to 4 + one end
    a = golap.HostColumn.encode(np.full(4096, 7, np.int32), codec=3, page_rows=4096, chunk_rows=4096).arrays()
    assert a["page_csize"][0] == 16 + 4 + 4
    # arithmetic progression -> delta bit-pack with bw = 0: header only
    a = golap.HostColumn.encode(np.arange(4096, dtype=np.int32) * 3 + 5, codec=2, page_rows=4096,
                                chunk_rows=4096).arrays()
    assert a["page_csize"][0] == 32
    assert a["comp"][int(a["page_off"][0]) + 20] == 0  # bw byte
    # sorted with steps 0/1 -> 1 bit per delta
    v = np.cumsum(rng(1).integers(0, 2, 4096)).astype(np.int32)
    a = golap.HostColumn.encode(v, codec=2, page_rows=4096, chunk_rows=4096).arrays()
    assert a["page_csize"][0] == 32 + 4 * ((4095 + 31) // 32)
    # AUTO chooses the smallest of the four
    sizes = {}
    for c in (1, 2, 3, 4):
        sizes[c] = int(golap.HostColumn.encode(v, codec=c, page_rows=4096, chunk_rows=4096).arrays()["page_csize"][0])
    auto = golap.HostColumn.encode(v, codec=255, page_rows=4096, chunk_rows=4096).arrays()
    assert int(auto["page_csize"][0]) == min(sizes.values())


def test_for_closed_forms(golap):
    # range exactly 2^9 - 1 -> 9-bit fields: header + ceil(4096*9/8) + 8 padding bytes
    v = (100 + rng(2).integers(0, 512, 4096)).astype(np.int32)
    v[0], v[1] = 100, 611
    a = golap.HostColumn.encode(v, codec=4, page_rows=4096, chunk_rows=4096).arrays()
    off = int(a["page_off"][0])
    assert a["page_codec"][0] & 0x0f == 4 and a["page_codec"][0] & 0x40  # point-readable
    assert a["page_csize"][0] == 32 + 4096 * 9 // 8 + 8
    assert a["comp"][off + 20] == 9 and int(a["comp"][off:off + 8].view(np.int64)[0]) == 100
    # constant -> 0-bit fields
    a = golap.HostColumn.encode(np.full(4096, -5, np.int16), codec=4, page_rows=4096, chunk_rows=4096).arrays()
    assert a["page_csize"][0] == 32 + 8
    # a range above 2^56 does not fit: the page falls back to LZ4
    w = np.array([0, 1 << 60] * 2048, np.int64)
    a = golap.HostColumn.encode(w, codec=4, page_rows=4096, chunk_rows=4096).arrays()
    assert a["page_codec"][0] & 0x0f in (1, 5)


def test_for_known_answer(orc, golap):
    """Fields [5, 0, 7, 1] of 3 bits over base 10, LSB first: bytes 0xC5 0x03 (hand-computed)."""
    page = np.zeros(32 + 2 + 8, np.uint8)
    page[0:8] = np.array([10], np.int64).view(np.uint8)
    page[16:20] = np.array([4], np.uint32).view(np.uint8)
    page[20], page[21] = 3, 4
    page[32], page[33] = 0xC5, 0x03
    r, out = orc.page_decode(4, page, 16)
    assert r == 4 and list(out[:16].view(np.int32)) == [15, 10, 17, 11]
    a = golap.HostColumn.encode(np.array([15, 10, 17, 11], np.int32), codec=4, page_rows=64,
                                chunk_rows=64).arrays()
    off = int(a["page_off"][0])
    assert bytes(a["comp"][off:off + a["page_csize"][0]]) == page.tobytes()
    # a truncated page is rejected
    assert orc.page_decode(4, page[:-1], 16)[0] == -1


def test_assemble_from_liblz4_pages(orc, golap, lz4lib):
    values = gen("text_like", 0, np.uint8) if False else np.frombuffer(
        (b"lorem ipsum dolor sit amet " * 5000)[:40000], dtype=np.uint8).copy()
    pages = [lz4lib.compress(values[i:i + 4096].tobytes()) for i in range(0, values.size, 4096)]
    hc = golap.HostColumn.assemble(2, values.size, 4096, 16384, pages, [1] * len(pages))
    a = hc.arrays()
    out = orc.decode_column(a)
    assert out[:values.size].tobytes() == values.tobytes()
    assert np.all(a["chunk_min"] == np.iinfo(np.int64).min)  # unknown statistics never prune


def test_encode_rejects_bad_options(golap):
    with pytest.raises(golap.GolapError):
        golap.HostColumn.encode(np.zeros(100, np.int32), page_rows=100)  # not a multiple of 64
    with pytest.raises(golap.GolapError):
        golap.HostColumn.encode(np.zeros(100, np.int32), page_rows=1024, chunk_rows=1536)


@pytest.mark.parametrize("min_match", [4, 5, 8])
def test_literal_only_flag_means_no_match(orc, golap, min_match):
    """A page is flagged literal-only (read in place by the scans) only if the block is one
    literal run: values < 2^23 give blocks with a single short match whose size equals the
    literal-only size, which must not be flagged."""
    r = rng(17)
    v = r.integers(100, 5_000_000, 2_000_000).astype(np.int32)
    a = golap.HostColumn.encode(v, codec=1, page_rows=4096, chunk_rows=65536, lz4_min_match=min_match).arrays()
    d = a["desc"]
    raw = v.view(np.uint8)
    flagged = 0
    for p in range(d.n_pages):
        rows = min(4096, d.n_rows - p * 4096)
        nb = rows * 4
        off, cs = int(a["page_off"][p]), int(a["page_csize"][p])
        if a["page_codec"][p] & PAGE_LITERAL_ONLY:
            flagged += 1
            hdr = 1 if nb < 15 else 2 + (nb - 15) // 255
            assert np.array_equal(a["comp"][off + hdr:off + hdr + nb], raw[p * 16384:p * 16384 + nb]), p
        blk = lz4_block(a["comp"][off:off + cs], a["page_codec"][p])
        assert orc.lz4_decode(blk, nb) == raw[p * 16384:p * 16384 + nb].tobytes()
    assert flagged > 0
