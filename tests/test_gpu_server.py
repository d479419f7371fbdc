"""Serving loop: requests admitted by HBM budget and continuously batched into persistent fused
launches on several streams. Every result must be the same alone (golap_scan_run /
golap_star_run) and in the mix, bit-exact."""
import ctypes
import threading

import numpy as np
import pytest

from gpu_util import say
from test_gpu_batch import Fixture, make_queries
from test_gpu_scan import scan_mix, template, trip_cols

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trip(golap, ctx):
    host, dev = trip_cols(golap, ctx, 3_000_000, seed=5)
    yield host, dev
    for d in dev:
        ctx.free_column(d)


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a[1:4], b[1:4]))


def scan_requests(dev, rng, n):
    out = []
    for t, h0, h1, full in scan_mix(rng, n):
        preds, keys, vc = template(t, h0, h1, full_hours=full)
        out.append(dict(cols=dev, preds=preds, keys=keys, value_col=vc))
    return out


def test_server_scan_results_alone_and_in_mix(golap, ctx, trip):
    host, dev = trip
    reqs = scan_requests(dev, np.random.default_rng(31), 48)
    alone = [ctx.scan_run(**p) for p in reqs]
    with ctx.server(slots=2, max_batch=16, emit_threads=4) as srv:
        got = [None] * len(reqs)

        def client(k):
            for i in range(k, len(reqs), 4):  # four client threads submit and wait concurrently
                t = srv.submit_scan(**reqs[i])
                got[i] = srv.wait(t)

        th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        st = srv.stats()
    say("server stats", st)
    for i in range(len(reqs)):
        assert same(got[i], alone[i]), i
        assert got[i][0].rows_matched == alone[i][0].rows_matched and got[i][0].chunks_run == alone[i][0].chunks_run
        assert got[i][4] > 0
    assert st["completed"] == st["submitted"] == len(reqs) and st["refused"] == 0


def test_server_burst_forms_batches(golap, ctx, trip):
    """A burst submitted at once is served in batches of several requests each."""
    host, dev = trip
    reqs = scan_requests(dev, np.random.default_rng(32), 40)
    with ctx.server(slots=2, max_batch=64) as srv:
        tickets = [srv.submit_scan(**p) for p in reqs]
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
    for p, g in zip(reqs, got):
        assert same(g, ctx.scan_run(**p))
    assert st["batches"] < len(reqs) and st["mean_batch"] > 1


def test_server_budget_admission(golap, ctx, trip):
    host, dev = trip
    reqs = scan_requests(dev, np.random.default_rng(33), 12)
    # a budget of about two average requests: admission keeps batches small, results unchanged
    with ctx.server(slots=2, hbm_budget=6 << 20) as srv:
        ok, tickets = [], []
        for p in reqs:
            try:
                tickets.append(srv.submit_scan(**p))
                ok.append(p)
            except golap.GolapError as e:
                assert e.status == -6  # GOLAP_EBUDGET: a full-year request does not fit alone
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["peak_admitted_bytes"] <= 6 << 20
    assert st["refused"] == len(reqs) - len(ok)
    for p, g in zip(ok, got):
        assert same(g, ctx.scan_run(**p))


def test_server_refuses_invalid_plans(golap, ctx, trip):
    host, dev = trip
    with ctx.server() as srv:
        with pytest.raises(golap.GolapError):
            srv.submit_scan(cols=dev, preds=[(9, 0, 1)], keys=[], value_col=-1)  # no column 9
        t = srv.submit_scan(cols=dev, preds=[(1, 2, 6)], keys=[(1, 1, 1, 6, 1)], value_col=3)
        r = srv.wait(t)
        assert r[0].n_groups == 5
        assert srv.stats()["refused"] == 1


def test_server_star_and_scan_mixed(golap, ctx, trip):
    """Star requests over the star fixture and scan requests over the trip table, interleaved."""
    host, dev = trip
    fx = Fixture(golap, ctx, 1_000_000, seed=8)
    try:
        qs = make_queries(np.random.default_rng(34), 10, fx.nch)
        flags = ctx.alloc(fx.nch * 2)
        star_alone = [fx.single(q, flags) for q in qs]
        ctx.free(flags)
        scans = scan_requests(dev, np.random.default_rng(35), 10)
        scan_alone = [ctx.scan_run(**p) for p in scans]
        with ctx.server(slots=3) as srv:
            tickets = []
            for i in range(10):
                tickets.append(("star", i, srv.submit_star(**fx.batch_entry(qs[i]))))
                tickets.append(("scan", i, srv.submit_scan(**scans[i])))
            for kind, i, t in tickets:
                g = srv.wait(t)
                assert same(g, star_alone[i] if kind == "star" else scan_alone[i]), (kind, i)
    finally:
        fx.free()


def test_server_star_batches_with_concurrent_clients(golap, ctx, trip):
    """Two slots serve multi-query star batches (each batch frees its shared-column filter on the
    completion thread) while four client threads keep submitting: the frees race with the
    scheduler's next launch unless the context's allocation registry and error string are
    locked. Every result equals its single-query run, and a second wait on a ticket is refused."""
    fx = Fixture(golap, ctx, 1_000_000, seed=9)
    try:
        qs = make_queries(np.random.default_rng(36), 40, fx.nch)
        flags = ctx.alloc(fx.nch * 2)
        alone = [fx.single(q, flags) for q in qs]
        ctx.free(flags)
        entries = [fx.batch_entry(q) for q in qs]
        with ctx.server(slots=2, max_batch=8, emit_threads=4) as srv:
            got = [None] * len(qs)

            def client(k):  # each client submits its share, then waits on it
                mine = [(i, srv.submit_star(**entries[i])) for i in range(k, len(qs), 4)]
                for i, t in mine:
                    got[i] = srv.wait(t)

            th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
            for t in th:
                t.start()
            for t in th:
                t.join()
            st = srv.stats()
            t = srv.submit_star(**entries[0])
            assert same(srv.wait(t), alone[0])
            with pytest.raises(golap.GolapError):
                srv.wait(t)
            # two threads waiting on one ticket through the C ABI: exactly one gets the result
            t = srv.submit_star(**entries[1])
            dom = srv.shape.pop(t)[1]
            rcs = []

            def raw_wait():
                r = golap.StarResult()
                k, su, cn = np.zeros(dom * 2, np.int32), np.zeros(dom, np.int64), np.zeros(dom, np.uint64)
                rcs.append(ctx.lib.golap_server_wait(srv.h, t, ctypes.byref(r), k.ctypes.data, su.ctypes.data,
                                                     cn.ctypes.data, dom, None))

            th = [threading.Thread(target=raw_wait) for _ in range(2)]
            for x in th:
                x.start()
            for x in th:
                x.join()
            assert sorted(rcs) == [golap.EINVAL, 0], rcs
        say("star server stats", st)
        for i in range(len(qs)):
            assert same(got[i], alone[i]), i
        assert st["completed"] == st["submitted"] == len(qs) and st["mean_batch"] > 1
    finally:
        fx.free()
