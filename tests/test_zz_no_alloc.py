"""golap_scan_run_noalloc (no counterpart in the reference: golap_scan_run for callers that
cannot allocate device memory) against golap_scan_run on the same plans, bit-exact. Two calls in
a row, the second over a smaller input (fewer chunks and a smaller key domain), so that contents
the first call left in the shared device array would show; the library's device allocation is
unchanged across the calls; a plan beyond the array's size is refused and the next call works."""
import numpy as np
import pytest

from gpu_util import say
from test_gpu_scan import template, trip_cols

pytestmark = pytest.mark.gpu

FIELDS = ("n_groups", "rows_matched", "total_sum", "chunks_run", "chunks_pruned", "comp_bytes")


@pytest.fixture(scope="module")
def trip(golap, ctx):
    host, dev = trip_cols(golap, ctx, 1_500_000, seed=17)
    yield host, dev
    for d in dev:
        ctx.free_column(d)


def both(ctx, dev, preds, keys, vc, **kw):
    want = ctx.scan_run(dev, preds, keys, vc, **kw)
    before = ctx.lib.golap_device_allocated(ctx.h)
    got = ctx.scan_run(dev, preds, keys, vc, noalloc=True, **kw)
    assert ctx.lib.golap_device_allocated(ctx.h) == before
    for f in FIELDS:
        assert getattr(got[0], f) == getattr(want[0], f), f
    for a, b in zip(got[1:], want[1:]):
        assert np.array_equal(a, b)
    return got[0]


@pytest.mark.parametrize("agg", [0, 2])
def test_noalloc_scan_equals_scan_run(golap, ctx, trip, agg):
    host, dev = trip
    # zone x hour over 900 hours (238,500 groups, about 3.8 MB of the array), then a later,
    # shorter range with a passenger predicate (fewer chunks, 26,765 groups)
    preds, keys, vc = template(0, 1000, 1899)
    say("noalloc scan: 900 hours, agg", agg)
    r1 = both(ctx, dev, preds, keys, vc, agg_mode=agg)
    assert r1.n_groups > 10_000 and r1.chunks_run > 0
    preds, keys, vc = template(1, 4000, 4100)
    say("noalloc scan: 101 hours, agg", agg)
    r2 = both(ctx, dev, preds, keys, vc, agg_mode=agg)
    assert 0 < r2.n_groups < r1.n_groups and r2.chunks_run < r1.chunks_run
    # a chunk sub-range without zone pruning (every chunk of it is a work item)
    say("noalloc scan: chunks 3..9, no zone pruning")
    both(ctx, dev, [(1, 2, 6)], [(1, 1, 1, 6, 1)], 3, chunk_begin=3, chunk_end=9, zone_prune=0, agg_mode=agg)


def test_noalloc_scan_refuses_plans_beyond_the_array(golap, ctx, trip):
    host, dev = trip
    # zone x hour over the whole year: 2,327,760 groups, 37 MB of dense result
    preds, keys, vc = template(0, 0, 8783, full_hours=True)
    say("noalloc scan: refused plan")
    with pytest.raises(golap.GolapError) as e:
        ctx.scan_run(dev, preds, keys, vc, noalloc=True)
    assert e.value.status == golap.ECAPACITY
    preds, keys, vc = template(2, 300, 400)
    say("noalloc scan: after the refusal")
    both(ctx, dev, preds, keys, vc)
