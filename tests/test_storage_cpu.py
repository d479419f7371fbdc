"""Column files (the DenseWriter / .dat serialization analogue, core/access.hpp:116-157,
core/hl/serialization.hpp:60-126): a saved column reads back identical, its page data starts
on a 4 KiB boundary, and damaged files are refused."""
import os

import numpy as np
import pytest

import golap_public_amd as g


@pytest.mark.parametrize("codec", [g.CODEC_RAW, g.CODEC_LZ4, g.CODEC_DBP, g.CODEC_RLE, g.CODEC_FOR, g.CODEC_AUTO])
def test_column_file_round_trip(tmp_path, codec):
    rng = np.random.default_rng(codec)
    v = np.sort(rng.integers(0, 1 << 20, 300_001)).astype(np.int32)
    hc = g.HostColumn.encode(v, codec=codec, page_rows=2048, chunk_rows=16384)
    path = tmp_path / "col.gcol"
    hc.save(path)
    back = g.HostColumn.load_file(path)
    a, b = hc.arrays(), back.arrays()
    for k in ("comp", "page_off", "page_csize", "page_codec", "chunk_min", "chunk_max"):
        assert np.array_equal(a[k], b[k]), k
    for f in ("n_rows", "n_pages", "n_chunks", "type", "elem", "page_rows", "chunk_rows", "payload_bytes"):
        assert getattr(a["desc"], f) == getattr(b["desc"], f), f
    size = os.path.getsize(path)
    assert size % 4096 == 0
    # the page data is the file's last region, 4 KiB aligned
    data_off = size - ((a["comp"].size + 4095) // 4096) * 4096
    assert data_off % 4096 == 0
    with open(path, "rb") as f:
        f.seek(data_off)
        assert f.read(a["comp"].size) == a["comp"].tobytes()
    hc.free()
    back.free()


def test_damaged_column_files_are_refused(tmp_path):
    hc = g.HostColumn.encode(np.arange(50_000, dtype=np.int32), codec=g.CODEC_LZ4, page_rows=4096, chunk_rows=16384)
    path = tmp_path / "col.gcol"
    hc.save(path)
    raw = bytearray(open(path, "rb").read())
    bad_magic = bytearray(raw)
    bad_magic[0] ^= 0xFF
    (tmp_path / "m.gcol").write_bytes(bytes(bad_magic))
    (tmp_path / "t.gcol").write_bytes(bytes(raw[: len(raw) - 8192]))  # truncated page data
    bad_rows = bytearray(raw)
    bad_rows[32:40] = (10**9).to_bytes(8, "little")  # n_rows no longer matches the page count
    (tmp_path / "r.gcol").write_bytes(bytes(bad_rows))
    for name in ("m.gcol", "t.gcol", "r.gcol", "missing.gcol"):
        with pytest.raises(g.GolapError):
            g.HostColumn.load_file(tmp_path / name)
    hc.free()
