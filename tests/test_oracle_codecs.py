"""The oracle's LZ4 block decoder against the reference's CPU codec (liblz4, the library that
core/comp_cpu.hpp:22-32 calls; lz4 @ce8ee024 in scripts/init.sh:32, 1.9.3 installed here) and
the round-trip patterns of core/test/comp_tests.cu:18-51."""
import numpy as np
import pytest

from conftest import rng


@pytest.mark.parametrize("size", [1 << 12, 1 << 15, 1 << 20])
def test_comp_tests_pattern_roundtrip(orc, lz4lib, size):
    data = bytes(((size - i) & 0xFF) for i in range(size))  # (uint8)(size-i), comp_tests.cu:27
    block = lz4lib.compress(data)
    assert orc.lz4_decode(block, size) == data
    assert lz4lib.decompress_partial(block, size) == data


@pytest.mark.parametrize("kind", ["random", "lowcard", "runs", "text", "empty", "tiny"])
def test_liblz4_blocks_decode_identically(orc, lz4lib, kind):
    r = rng(7)
    if kind == "random":
        data = r.integers(0, 256, 70000, dtype=np.uint8).tobytes()
    elif kind == "lowcard":
        data = r.integers(0, 11, 65536, dtype=np.uint8).tobytes()
    elif kind == "runs":
        data = np.repeat(r.integers(0, 5, 400, dtype=np.uint8), r.integers(1, 300, 400)).tobytes()
    elif kind == "text":
        data = (b"the quick brown fox jumps over the lazy dog " * 2000)[:50000]
    elif kind == "empty":
        data = b""
    else:
        data = b"abc"
    block = lz4lib.compress(data)
    assert orc.lz4_decode(block, len(data)) == data
    assert lz4lib.decompress_partial(block, len(data)) == data


def test_overlapping_match_kat(orc, lz4lib):
    # literal 'a', match offset 1 length 19 (period-1 overlap), then 5 literals
    block = bytes([0x1F, ord("a"), 0x01, 0x00, 0x00, 0x50]) + b"bcdef"
    want = b"a" * 20 + b"bcdef"
    assert lz4lib.decompress_partial(block, len(want)) == want
    assert orc.lz4_decode(block, len(want)) == want
    # period-3 overlap: 'xyz' then match offset 3 length 30
    block = bytes([0x3F, ord("x"), ord("y"), ord("z"), 0x03, 0x00, 0x0B, 0x50]) + b"12345"
    want = b"xyz" * 11 + b"12345"
    assert lz4lib.decompress_partial(block, len(want)) == want
    assert orc.lz4_decode(block, len(want)) == want


@pytest.mark.parametrize("bad", ["truncated_literals", "offset_zero", "offset_before_start",
                                 "truncated_offset", "overlong_output"])
def test_malformed_blocks_rejected_like_liblz4(orc, lz4lib, bad):
    good = bytes([0x1F, ord("a"), 0x01, 0x00, 0x00, 0x50]) + b"bcdef"
    n = 25
    if bad == "truncated_literals":
        block = good[:-2]
    elif bad == "offset_zero":
        block = bytes([0x1F, ord("a"), 0x00, 0x00, 0x00, 0x50]) + b"bcdef"
    elif bad == "offset_before_start":
        block = bytes([0x1F, ord("a"), 0x05, 0x00, 0x00, 0x50]) + b"bcdef"
    elif bad == "truncated_offset":
        block = bytes([0x1F, ord("a"), 0x01])
    else:
        block = good
        n = 10  # output larger than the destination
    assert orc.lz4_decode(block, n) is None
    # liblz4 1.9.3's safe_partial stops at the target size on an overlong block, returns the
    # bytes decoded so far when the input ends inside a sequence header, and accepts a zero
    # offset (the block format forbids it); the oracle rejects all three. In every case the
    # library does not produce the full, correct output either.
    lib = lz4lib.decompress_partial(block, n)
    if bad in ("truncated_literals", "offset_before_start"):
        assert lib is None
    elif bad == "truncated_offset":
        assert lib is None or len(lib) < n
