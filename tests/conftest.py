import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def golap():
    import golap_public_amd
    golap_public_amd.load()
    return golap_public_amd


@pytest.fixture(scope="session")
def ctx(golap):
    """One device context for the whole GPU session."""
    c = golap.Context(0)
    yield c
    c.close()


class Lz4Lib:
    """The reference's CPU codec (liblz4, core/comp_cpu.hpp:22-32), called as a library."""

    def __init__(self):
        self.lib = ctypes.CDLL("liblz4.so.1")
        cp, ci = ctypes.c_char_p, ctypes.c_int
        self.lib.LZ4_compress_default.argtypes = [cp, cp, ci, ci]
        self.lib.LZ4_compress_default.restype = ci
        self.lib.LZ4_decompress_safe.argtypes = [cp, cp, ci, ci]
        self.lib.LZ4_decompress_safe.restype = ci
        self.lib.LZ4_decompress_safe_partial.argtypes = [cp, cp, ci, ci, ci]
        self.lib.LZ4_decompress_safe_partial.restype = ci
        self.lib.LZ4_compressBound.argtypes = [ci]
        self.lib.LZ4_compressBound.restype = ci

    def compress(self, data: bytes) -> bytes:
        cap = self.lib.LZ4_compressBound(len(data))
        out = ctypes.create_string_buffer(cap)
        n = self.lib.LZ4_compress_default(data, out, len(data), cap)
        assert n > 0
        return out.raw[:n]

    def decompress_partial(self, block: bytes, n_out: int):
        """LZ4_decompress_safe_partial(in, out, comp, n, n) exactly as comp_cpu.hpp:26 calls it."""
        out = ctypes.create_string_buffer(max(n_out, 1))
        r = self.lib.LZ4_decompress_safe_partial(block, out, len(block), n_out, n_out)
        return None if r < 0 else out.raw[:r]


@pytest.fixture(scope="session")
def lz4lib():
    try:
        return Lz4Lib()
    except OSError:
        pytest.skip("liblz4.so.1 not present")


def rng(seed=1234):
    return np.random.default_rng(seed)


LZ4I = 5


def lz4i_header_bytes(page):
    """Header bytes of an LZ4I page: {u16 nseq, u8 seg_log2, u8 unit_log2}, u16 offsets of every
    2^seg_log2-th sequence's token, zero padding to a multiple of 16."""
    page = bytes(page)
    nseq, sl = page[0] | (page[1] << 8), page[2]
    nseg = -(-nseq // (1 << sl))
    return (4 + 2 * (nseg - 1) + 15) // 16 * 16


def lz4_block(page, codec):
    """The standard LZ4 block of an LZ4 (codec 1) or LZ4I (codec 5) page."""
    page = bytes(page)
    return page[lz4i_header_bytes(page):] if codec & 0x0F == LZ4I else page
