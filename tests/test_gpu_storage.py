"""SSD -> HBM: a column file read through a ring of pinned buffers (O_DIRECT preads overlapped
with hipMemcpyAsync, the SSD2CPU2GPU path of core/apps.cuh:315-320) gives a device column whose
bytes and plan results equal the uploaded column's."""
import os

import numpy as np
import pytest

from gpu_util import q1_dataset, say
from test_gpu_q1 import date_map, release, upload_q1

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("segment,buffers", [(4096, 2), (65536, 3), (0, 0)])
def test_loaded_q1_columns_equal_uploaded(orc, golap, ctx, tmp_path, segment, buffers):
    ds = q1_dataset(orc, 1_500_000, seed=44)
    cols, hosts = upload_q1(ctx, golap, ds, codec=1, page_rows=4096, chunk_rows=65536, lz4_min_match=16)
    loaded = []
    m, bufs = date_map(ctx, golap, ds, 1995)
    try:
        for i, h in enumerate(hosts):
            path = tmp_path / f"c{i}.gcol"
            h.save(path)
            col, st = ctx.load_file(path, segment, buffers)
            loaded.append(col)
            say("load", i, st.as_dict())
            a = h.arrays()
            d = ctx.column_desc(col)
            assert st.bytes == a["comp"].size and d.comp_bytes == a["comp"].size
            assert np.array_equal(ctx.to_host(d.comp, a["comp"].size, np.uint8), a["comp"])
            assert np.array_equal(ctx.to_host(d.page_off, d.n_pages, np.uint64), a["page_off"])
        want = ctx.q1_run(*cols, m, (4, 6), (30, 50))
        got = ctx.q1_run(*loaded, m, (4, 6), (30, 50))
        assert (got.sum, got.matches, got.comp_bytes) == (want.sum, want.matches, want.comp_bytes)
    finally:
        for c in loaded:
            ctx.free_column(c)
        release(ctx, cols, m, bufs)


def test_load_throughput_and_refusals(golap, ctx, tmp_path):
    n = 100_000_000  # 400 MB of raw int32 page data
    v = np.random.default_rng(1).integers(0, 1 << 30, n).astype(np.int32)
    h = golap.HostColumn.encode(v, codec=golap.CODEC_RAW, page_rows=4096, chunk_rows=262144)
    path = tmp_path / "big.gcol"
    h.save(path)
    col, st = ctx.load_file(path, 64 << 20, 4)
    try:
        gbps = st.bytes / (st.total_ms * 1e-3) / 1e9
        say("load", st.as_dict(), f"{gbps:.2f} GB/s end to end")
        d = ctx.column_desc(col)
        back = ctx.to_host(d.comp, d.comp_bytes, np.uint8)
        assert np.array_equal(back, h.arrays()["comp"])
    finally:
        ctx.free_column(col)
    with open(path, "r+b") as f:  # a damaged header is refused before any device work
        f.write(b"X")
    with pytest.raises(golap.GolapError):
        ctx.load_file(path)
    with pytest.raises(golap.GolapError):
        ctx.load_file(tmp_path / "missing.gcol")
