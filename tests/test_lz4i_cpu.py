"""LZ4I pages (an LZ4 block behind a parse index, this build's page format) on the CPU.

What pins them:
- the block is standard LZ4: liblz4 (the reference's CPU codec, core/comp_cpu.hpp:22-32) decodes
  it to the page's values;
- the header is checked against an independent parse of the block written here in Python
  (token positions, sequence count), not against the encoder's own bookkeeping;
- the oracle (oracle_lz4i_decode) accepts exactly the pages whose index agrees with that parse:
  damaged entries, counts, padding or segment lengths are refused, and a page re-indexed here
  with another segment length is still accepted (the format does not depend on the encoder's
  choice);
- typed pages (2- and 4-byte columns, the reference's typed nvcomp LZ4, core/apps.cuh:262-270)
  record their unit in header byte 3, and the Python parse here checks that every literal run,
  offset and match is a whole number of units; the oracle refuses a unit the block does not keep."""
import numpy as np
import pytest

from conftest import LZ4I, lz4_block, lz4i_header_bytes, rng


def parse_tokens(blk):
    """Token offsets of an LZ4 block's sequences (lz4_Block_format.md), or None."""
    ip, toks = 0, []
    while True:
        if ip >= len(blk):
            return None
        toks.append(ip)
        t = blk[ip]
        ip += 1
        lit = t >> 4
        if lit == 15:
            while True:
                if ip >= len(blk):
                    return None
                b = blk[ip]
                ip += 1
                lit += b
                if b != 255:
                    break
        ip += lit
        if ip > len(blk):
            return None
        if ip == len(blk):
            return toks
        ip += 2
        if ip > len(blk):
            return None
        if t & 15 == 15:
            while True:
                if ip >= len(blk):
                    return None
                b = blk[ip]
                ip += 1
                if b != 255:
                    break


def parse_sequences(blk):
    """(literal length, offset, match length) of every sequence of a well-formed LZ4 block; the
    last sequence has offset and match length 0."""
    ip, out = 0, []
    while True:
        t = blk[ip]
        ip += 1
        lit = t >> 4
        if lit == 15:
            while True:
                b = blk[ip]
                ip += 1
                lit += b
                if b != 255:
                    break
        ip += lit
        if ip == len(blk):
            out.append((lit, 0, 0))
            return out
        off = blk[ip] | (blk[ip + 1] << 8)
        ip += 2
        ml = t & 15
        if ml == 15:
            while True:
                b = blk[ip]
                ip += 1
                ml += b
                if b != 255:
                    break
        out.append((lit, off, ml + 4))


def make_page(blk, sl, unit_log2=0):
    """An LZ4I page of block blk with segments of 2^sl sequences, claiming units of 2^unit_log2
    bytes."""
    toks = parse_tokens(blk)
    n = len(toks)
    nseg = -(-n // (1 << sl))
    hb = (4 + 2 * (nseg - 1) + 15) // 16 * 16
    h = bytearray(hb)
    h[0], h[1], h[2], h[3] = n & 0xFF, n >> 8, sl, unit_log2
    for s in range(1, nseg):
        t = toks[s << sl]
        h[4 + 2 * (s - 1)], h[5 + 2 * (s - 1)] = t & 0xFF, t >> 8
    return bytes(h) + bytes(blk)


def encoder_sl(n):
    """The encoder's segment length: the smallest power of two >= 8 giving <= 256 segments."""
    sl = 3
    while (256 << sl) < n:
        sl += 1
    return sl


def columns():
    r = rng(41)
    n = 4096 * 6 + 123
    dates = (19920101 + r.integers(0, 2556, n)).astype(np.int32)  # many 4-byte matches
    yield "date", dates
    yield "lowcard_u8", r.integers(0, 11, n).astype(np.uint8)
    yield "runs_i16", np.repeat(r.integers(0, 5, n // 40 + 1), 40)[:n].astype(np.int16)
    yield "sorted_i64", np.sort(r.integers(0, 1 << 40, n)).astype(np.int64)


@pytest.mark.parametrize("name,values", list(columns()), ids=[c[0] for c in columns()])
@pytest.mark.parametrize("page_rows", [1024, 4096])
@pytest.mark.parametrize("typed", [False, True])
def test_encoder_lz4i_pages_agree_with_parse_and_liblz4(orc, golap, lz4lib, name, values, page_rows, typed):
    hc = golap.HostColumn.encode(values, codec=1, page_rows=page_rows, chunk_rows=page_rows * 4, lz4_min_match=4,
                                 lz4_unit=0 if typed else 1)
    a = hc.arrays()
    d = a["desc"]
    raw = values.view(np.uint8)
    n_lz4i = 0
    for p in range(d.n_pages):
        rows = min(page_rows, d.n_rows - p * page_rows)
        nb = rows * values.itemsize
        off, cs, code = int(a["page_off"][p]), int(a["page_csize"][p]), int(a["page_codec"][p])
        page = a["comp"][off:off + cs].tobytes()
        want = raw[p * page_rows * values.itemsize:p * page_rows * values.itemsize + nb].tobytes()
        if code & 0x0F != LZ4I:
            assert code & 0x0F == 1 and code & 0x80  # only one-literal-run blocks stay plain
            continue
        n_lz4i += 1
        assert off % 16 == 0
        blk = lz4_block(page, code)
        toks = parse_tokens(blk)
        assert toks is not None
        nseq, sl = page[0] | (page[1] << 8), page[2]
        unit = min(values.itemsize, 4) if typed else 1
        assert nseq == len(toks) and sl == encoder_sl(nseq) and page[3] == unit.bit_length() - 1
        assert all(lit % unit == 0 and off % unit == 0 and ml % unit == 0 for lit, off, ml in parse_sequences(blk))
        hb = lz4i_header_bytes(page)
        nseg = -(-nseq // (1 << sl))
        assert nseg <= 256 and hb % 16 == 0
        for s in range(1, nseg):
            assert page[4 + 2 * (s - 1)] | (page[5 + 2 * (s - 1)] << 8) == toks[s << sl]
        assert not any(page[4 + 2 * (nseg - 1):hb])
        assert lz4lib.decompress_partial(blk, nb) == want
        r, out = orc.page_decode(code, page, nb)
        assert r == nb and out[:nb].tobytes() == want
    if name in ("date", "lowcard_u8", "runs_i16"):
        assert n_lz4i >= d.n_pages - 1  # the short last page may be one literal run


def test_oracle_refuses_damaged_indices(orc, golap):
    r = rng(42)
    v = (19920101 + r.integers(0, 2556, 4096)).astype(np.int32)
    a = golap.HostColumn.encode(v, codec=1, page_rows=4096, chunk_rows=4096, lz4_min_match=4, lz4_unit=0).arrays()
    page = a["comp"][int(a["page_off"][0]):int(a["page_off"][0]) + int(a["page_csize"][0])].tobytes()
    assert a["page_codec"][0] == LZ4I
    want = v.tobytes()
    assert orc.page_decode(LZ4I, page, len(want))[0] == len(want)
    blk = lz4_block(page, LZ4I)
    nseq = page[0] | (page[1] << 8)
    assert nseq > 1000

    def dec(pg):
        return orc.page_decode(LZ4I, pg, len(want))[0]

    b = bytearray(page)
    b[4] ^= 1  # the second segment starts one byte off
    assert dec(bytes(b)) == -1
    b = bytearray(page)
    b[0] = (nseq + 1) & 0xFF  # one sequence too many
    b[1] = (nseq + 1) >> 8
    assert dec(bytes(b)) == -1
    b = bytearray(page)
    b[0], b[1] = (nseq - 1) & 0xFF, (nseq - 1) >> 8  # one too few
    assert dec(bytes(b)) == -1
    b = bytearray(page)
    b[2] += 1  # longer segments: the entries point at the wrong sequences
    assert dec(bytes(b)) == -1
    assert page[3] == 2  # 4-byte units: the column's element width
    b = bytearray(page)
    b[3] = 3  # no such unit
    assert dec(bytes(b)) == -1
    b = bytearray(page)
    b[3] = 0  # a unit-aligned block may claim nothing
    assert dec(bytes(b)) == len(want)
    # a byte-granular block of the same values does not keep 4-byte units: claiming them is refused
    a1 = golap.HostColumn.encode(v, codec=1, page_rows=4096, chunk_rows=4096, lz4_min_match=4, lz4_unit=1).arrays()
    p1 = a1["comp"][int(a1["page_off"][0]):int(a1["page_off"][0]) + int(a1["page_csize"][0])].tobytes()
    assert p1[3] == 0 and orc.page_decode(LZ4I, p1, len(want))[0] == len(want)
    assert any(lit % 4 or off % 4 or ml % 4 for lit, off, ml in parse_sequences(lz4_block(p1, LZ4I)))
    b = bytearray(p1)
    b[3] = 2
    assert dec(bytes(b)) == -1
    hb = lz4i_header_bytes(page)
    nseg = -(-nseq // (1 << page[2]))
    if 4 + 2 * (nseg - 1) < hb:
        b = bytearray(page)
        b[hb - 1] = 7  # padding must be zero
        assert dec(bytes(b)) == -1
    # the same block indexed with other segment lengths is a valid page
    for sl in (4, 6, 9):
        pg = make_page(blk, sl, 2)
        assert dec(pg) == len(want)
    # a block that no longer parses is refused whatever the index says
    assert dec(make_page(blk, 5)[:-1]) == -1


@pytest.mark.parametrize("use_liblz4", [False, True])
def test_decode_timing_checksum_matches_values(orc, golap, use_liblz4):
    """The CPU baseline's decode lines (oracle decoders, liblz4) decode every page of a Q1-shaped
    table to its values: their checksum equals one computed from the raw columns."""
    r = rng(43)
    n = 4096 * 5 + 77
    cols = [(19920101 + r.integers(0, 2556, n)).astype(np.int32), r.integers(0, 11, n).astype(np.uint8),
            r.integers(1, 51, n).astype(np.uint8), r.integers(90000, 10500001, n).astype(np.int32)]
    arrays = [golap.HostColumn.encode(c, codec=1, page_rows=4096, chunk_rows=8192, lz4_min_match=4).arrays()
              for c in cols]
    assert (arrays[0]["page_codec"] == LZ4I).sum() >= 5
    st, cs, sec = orc.decode_timing(arrays, [c.itemsize for c in cols], n, 4096, use_liblz4, threads=3)
    if st == -2:
        pytest.skip("liblz4 missing")
    assert st == 0 and sec >= 0
    assert cs == orc.decode_checksum(cols)
    bad = [dict(a) for a in arrays]
    bad[0]["comp"] = bad[0]["comp"].copy()
    mid = int(bad[0]["page_off"][1]) + int(bad[0]["page_csize"][1]) // 2
    bad[0]["comp"][mid] ^= 0x5A  # a byte in the middle of page 1's block (behind its index)
    st2, cs2, _ = orc.decode_timing(bad, [c.itemsize for c in cols], n, 4096, use_liblz4, threads=3)
    assert st2 == -1 or cs2 != cs
