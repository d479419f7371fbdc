"""Writes tests/golden/liblz4_pages.npz: pages compressed by liblz4's LZ4_compress_default (the
reference's CPU codec, core/comp_cpu.hpp:30) and the bytes they decode to. The GPU decoder must
reproduce the original bytes from these third-party-made blocks.

Run: python tests/golden/make_lz4_fixture.py
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    lib = ctypes.CDLL("liblz4.so.1")
    lib.LZ4_compress_default.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    lib.LZ4_compress_default.restype = ctypes.c_int
    r = np.random.default_rng(20240601)
    parts = [
        np.frombuffer((b"SELECT sum(lo_extendedprice*lo_discount) FROM lineorder " * 400)[:16384], np.uint8),
        r.integers(0, 11, 16384).astype(np.uint8),                      # discount-like
        np.repeat(r.integers(0, 256, 128), 128).astype(np.uint8),       # runs
        r.integers(0, 256, 16384).astype(np.uint8),                     # incompressible
        np.tile(np.arange(3, dtype=np.uint8), 5462)[:16384],            # period-3 overlap
        np.zeros(16384, np.uint8),                                      # period-1 overlap
    ]
    data = np.concatenate(parts)
    data = np.concatenate([data, r.integers(0, 4, 1000).astype(np.uint8)])  # ragged last page
    page = 4096
    blocks, sizes = [], []
    for i in range(0, data.size, page):
        src = data[i:i + page].tobytes()
        out = ctypes.create_string_buffer(len(src) + len(src) // 255 + 16)
        n = lib.LZ4_compress_default(src, out, len(src), len(out))
        assert n > 0
        blocks.append(np.frombuffer(out.raw[:n], np.uint8))
        sizes.append(n)
    np.savez_compressed(os.path.join(HERE, "liblz4_pages.npz"), data=data, page_rows=page,
                        sizes=np.array(sizes, np.uint32), blob=np.concatenate(blocks))


if __name__ == "__main__":
    main()
