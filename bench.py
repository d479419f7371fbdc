"""Benchmark: the fused compressed-chunk -> filter -> join-probe -> int64-sum pipeline on one GPU.

Workload (one "step" = one query): 600,000,000 fact rows (date_key int32 over 2,556 days,
discount uint8 U[0,10], quantity uint8 U[1,50], price_cents int32 U[90000,10500000]) stored as
LZ4 pages (minimum match 4, as LZ4_compress_default codes them; pages with matches carry a parse
index) in 262,144-row chunks, resident in HBM. Each step builds the date hash table
(year == 1995) on the device and runs the fused kernel for
discount in [4,6] and quantity >= 30, SUM(price*discount) as int64.

Prints one JSON line: rows/s over the whole job, the roofline of the fused kernel (algorithmic
bytes / kernel time, against 8 TB/s), the CPU path (oracle decode + plan on the host cores, and
decode alone by the oracle and by liblz4) timed on the same pages, and as secondary fields the
same plan over minimum-match-16 pages and a serving sweep of the trip-table workload (served
and plain queries/s and p50/p99 at 0.5x and 0.9x the plain capacity and 0.9x the served one).
Other workloads: --workload star | batch | serve | q1_files | encode.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "fact rows/s and % HBM roofline (fused pipeline); queries/s + p99 latency, 1 GPU"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["q1", "star", "batch", "serve", "q1_files", "encode"], default="q1",
                    help="q1: fused filter/probe/sum (the headline line); star: star join + group-by; "
                         "q1_files: the q1 plan streamed from column files; encode: the q1 columns coded "
                         "on the device")
    ap.add_argument("--stream-dir", default=None,
                    help="q1_files workload: folder for the column files (default: a temporary folder)")
    ap.add_argument("--queries", type=int, default=64, help="batch workload: queries per batch")
    ap.add_argument("--zone-rows", type=int, default=65536, help="batch workload: rows per zone-map chunk")
    ap.add_argument("--trip-rows", type=int, default=1_500_000_000, help="serve workload: trip-table rows")
    ap.add_argument("--requests", type=int, default=1500, help="serve workload: arrivals per rate")
    ap.add_argument("--sweep-requests", type=int, default=1500,
                    help="q1 workload: requests per rate of the serving sweep reported beside it")
    ap.add_argument("--no-serve-sweep", action="store_true", help="q1 workload: skip the serving sweep")
    ap.add_argument("--no-mm16", action="store_true", help="q1 workload: skip the min-match 16 run")
    ap.add_argument("--budget-mb", type=int, default=4096, help="serve workload: HBM admission budget")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=600_000_000)
    ap.add_argument("--page-rows", type=int, default=4096)
    ap.add_argument("--chunk-rows", type=int, default=262144)
    ap.add_argument("--min-match", type=int, default=4,
                    help="LZ4 minimum match (4: what LZ4_compress_default emits)")
    ap.add_argument("--lz4-unit", type=int, default=1,
                    help="q1 workload: LZ4 sequence unit; 1: bytes (LZ4_compress_default), 0: whole "
                         "elements of the column's width (nvcomp's LZ4 manager built with the column's "
                         "data type, core/comp.cuh:230-236)")
    ap.add_argument("--seed", type=int, default=20240601)
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="after the timed steps, run 3 device copies of this many bytes (counter scale)")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "q1_fused_600M_traffic.json"))
    return ap.parse_args()


def cpu_baseline(args, hosts, key, year, gpu_sum):
    """The CPU path (CPU chunk decode + Q1 plan with thread-local partials) on the host cores."""
    import ctypes as C
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    lib = oracle.load()
    threads = len(os.sched_getaffinity(0))
    refs = (oracle.ColRef * 4)()
    for i, h in enumerate(hosts):
        d = h.desc()
        refs[i] = oracle.ColRef(d.comp, d.page_off, d.page_csize, d.page_codec)
    d0 = hosts[0].desc()
    preds, keep = oracle.make_preds([(year, 1995, 1995)])
    c = oracle.Q1Consts(4, 6, 30, 50, -2**63, 2**63 - 1)
    key = np.ascontiguousarray(key)
    times, res = [], None
    lib.oracle_q1_compressed(key.ctypes.data, key.size, C.cast(preds, C.c_void_p), 1, C.cast(refs, C.c_void_p),
                             d0.n_rows, d0.page_rows, d0.chunk_rows, 0, d0.n_chunks, C.byref(c), threads,
                             C.byref(C.c_uint64()))  # warm-up
    for _ in range(args.cpu_runs):
        m = C.c_uint64()
        t0 = time.perf_counter()
        s = lib.oracle_q1_compressed(key.ctypes.data, key.size, C.cast(preds, C.c_void_p), 1,
                                     C.cast(refs, C.c_void_p), d0.n_rows, d0.page_rows, d0.chunk_rows, 0,
                                     d0.n_chunks, C.byref(c), threads, C.byref(m))
        times.append(time.perf_counter() - t0)
        res = s
    med = statistics.median(times)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except Exception:
        pass
    log(f"cpu baseline: {threads} threads, median {med*1e3:.1f} ms, sum {res} (gpu {gpu_sum})")
    # decode alone, by the oracle's decoders and by liblz4 (the reference's CPU codec), on the
    # same pages: the checksums of the decoded bytes must agree
    addrs = [(h.desc().comp, h.desc().page_off, h.desc().page_csize, h.desc().page_codec) for h in hosts]
    elems = [4, 1, 1, 4]
    raw = d0.n_rows * sum(elems)
    decode = {}
    for name, use_lib in (("oracle", False), ("liblz4", True)):
        runs = []
        for _ in range(3):
            st, cs, sec = oracle.decode_timing(addrs, elems, d0.n_rows, d0.page_rows, use_lib, threads)
            if st != 0:
                break
            runs.append((sec, cs))
        if runs:
            sec = statistics.median(r[0] for r in runs)
            decode[name] = {"rows_per_s": d0.n_rows / sec, "decoded_GBps": raw / sec / 1e9, "ms": sec * 1e3,
                            "checksum": runs[-1][1]}
            log(f"cpu decode ({name}): {sec*1e3:.1f} ms, {raw/sec/1e9:.2f} GB/s decoded")
        else:
            decode[name] = {"status": st}
    if "checksum" in decode.get("oracle", {}) and "checksum" in decode.get("liblz4", {}):
        decode["checksums_equal"] = decode["oracle"]["checksum"] == decode["liblz4"]["checksum"]
    return {"value": d0.n_rows / med, "unit": "rows/s", "cores": threads, "kind": "port",
            "sample": f"full workload ({d0.n_rows} rows, same compressed pages), median of "
                      f"{args.cpu_runs} runs, CPU: {model}",
            "bit_exact_vs_gpu": bool(res == gpu_sum), "decode_only_median_of_3": decode}


def run_star(args):
    """Star join + group-by on 600M fact rows: date_key sorted (bit-packed delta), customer,
    supplier, part keys uniform (LZ4), revenue (LZ4); dimensions of 2,556 / 2,000,000 /
    200,000 / 1,400,000 rows; supplier region = 1 (1/5), customer nation = 7 (1/25), part
    brand = 42 (1/1000); group by (year, brand), SUM(revenue) int64 and COUNT."""
    import golap_public_amd as g
    n_cust, n_supp, n_part = 2_000_000, 200_000, 1_400_000
    key, year, _ = g.gen_date_dim(2556)
    log(f"generating {args.rows} star rows")
    fact = g.gen_star_fact(args.seed, args.rows, args.rows, key, n_cust, n_supp, n_part)
    nation = g.gen_uniform_i32(args.seed, 20, n_cust, 0, 24)
    region = g.gen_uniform_i32(args.seed, 21, n_supp, 0, 4)
    brand = g.gen_uniform_i32(args.seed, 22, n_part, 1, 1000)
    log("encoding")
    enc = dict(page_rows=args.page_rows, chunk_rows=args.chunk_rows, lz4_min_match=args.min_match)
    hosts = [g.HostColumn.encode(fact[0], codec=g.CODEC_DBP, **enc)] + \
            [g.HostColumn.encode(c, codec=g.CODEC_LZ4, **enc) for c in fact[1:]]
    del fact
    ctx = g.Context(0)
    cols = [ctx.upload(h) for h in hosts]
    comp = sum(h.desc().payload_bytes for h in hosts)
    log(f"compressed payload {comp/1e9:.3f} GB for {args.rows*20/1e9:.3f} GB raw")
    dev = lambda a: ctx.to_device(np.ascontiguousarray(a))
    d_key, d_year = dev(key), dev(year.astype(np.int32))
    c_key, c_nat = dev(np.arange(n_cust, dtype=np.int32)), dev(nation)
    s_key, s_reg = dev(np.arange(n_supp, dtype=np.int32)), dev(region)
    p_key, p_brand = dev(np.arange(n_part, dtype=np.int32)), dev(brand)
    maps = [ctx.hashmap(2556, d_key, g.I32), ctx.hashmap(n_cust, c_key, g.I32),
            ctx.hashmap(n_supp, s_key, g.I32), ctx.hashmap(n_part, p_key, g.I32)]

    def step():
        ctx.hashmap_build(maps[0], 2556, [], d_year)
        ctx.hashmap_build(maps[1], n_cust, [(c_nat, g.I32, 7, 7)])
        ctx.hashmap_build(maps[2], n_supp, [(s_reg, g.I32, 1, 1)])
        ctx.hashmap_build(maps[3], n_part, [(p_brand, g.I32, 42, 42)], p_brand)
        dims = [(cols[0], maps[0], 1992, 7), (cols[1], maps[1], 0, 0), (cols[2], maps[2], 0, 0),
                (cols[4 - 1], maps[3], 1, 1000)]
        return ctx.star_run(dims, cols[4])

    for _ in range(args.warmup):
        r = step()
    ctx.sync()
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = step()
        kms.append(r[0].kernel_ms)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    res, keys, sums, counts = r
    kernel_ms = statistics.mean(kms)
    log(f"groups {res.n_groups} rows {res.rows_matched} sum {res.total_sum}; "
        f"{elapsed*1e3/args.steps:.3f} ms/step, kernel {kernel_ms:.3f} ms")
    out = {"metric": "fact rows/s (star join + group-by, 1 GPU)", "value": args.rows * args.steps / elapsed,
           "unit": "rows/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "int64", "data": "synthetic: SplitMix64-seeded star schema",
           "config": {"workload": "star_groupby_600M", "rows": args.rows, "page_rows": args.page_rows,
                      "chunk_rows": args.chunk_rows, "codecs": "date_key DBP, keys/revenue LZ4",
                      "comp_bytes": res.comp_bytes, "kernel_ms": kernel_ms, "groups": res.n_groups,
                      "rows_matched": res.rows_matched, "result_sum": res.total_sum,
                      "driver_dim": res.driver_dim, "lds_group_slots": res.lds_group_slots,
                      "comp_bytes_per_s_GBps": res.comp_bytes / (kernel_ms * 1e-3) / 1e9},
           "roofline": {"bound": "hbm", "achieved": res.comp_bytes / (kernel_ms * 1e-3) / 1e9, "peak": 8000.0,
                        "unit": "GB/s", "frac": res.comp_bytes / (kernel_ms * 1e-3) / 1e9 / 8000.0,
                        "traffic": traffic_of("star_groupby_600M")}}
    # co-execution: the customer dimension joined on the host (the query 3 variants that keep
    # the customer table on the CPU), same groups as the all-device plan
    hm = g.HostHashMap(np.arange(n_cust, dtype=np.int32))
    hm.build([(nation, 7, 7)])
    dims = [(cols[0], maps[0], 1992, 7), (cols[1], None, 0, 0), (cols[2], maps[2], 0, 0), (cols[3], maps[3], 1, 1000)]
    cx = [ctx.star_run_hostdim(dims, cols[4], 1, hm) for _ in range(3)]
    same = all(np.array_equal(a, b) for a, b in zip(cx[-1][1:4], (keys, sums, counts)))
    cst = cx[-1][4]
    log(f"co-execution (customer on the host): {statistics.median(c[4].total_ms for c in cx):.1f} ms, "
        f"{cst.rounds} rounds, {cst.emitted} rows handed over, host {cst.host_ms:.1f} ms, same groups {same}")
    if not same:
        raise SystemExit("star: co-execution groups differ from the device plan's")
    out["config"]["coexec_customer_on_host"] = {
        "ms": statistics.median(c[4].total_ms for c in cx), "device_span_ms": cst.device_ms,
        "host_ms": cst.host_ms, "rounds": cst.rounds, "rows_to_host": cst.emitted, "same_groups": same}
    hm.free()
    calibrate(ctx, args)
    if not args.no_cpu_baseline:
        # the CPU plan of 02inmem.hpp over every fact row (decoded columns regenerated from the
        # same seed), timed and compared group by group with the device result
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = len(os.sched_getaffinity(0))
        f2 = oracle.gen_star_fact(args.seed, args.rows, args.rows, key, n_cust, n_supp, n_part)
        odims = [(key, [], year.astype(np.int32)), (np.arange(n_cust, dtype=np.int32), [(nation, 7, 7)], None),
                 (np.arange(n_supp, dtype=np.int32), [(region, 1, 1)], None),
                 (np.arange(n_part, dtype=np.int32), [(brand, 42, 42)], brand)]
        times, w = [], None
        for _ in range(max(1, min(args.cpu_runs, 3))):
            t0 = time.perf_counter()
            w = oracle.star_groupby(odims, list(f2[:4]), f2[4], threads=threads)
            times.append(time.perf_counter() - t0)
        del f2
        dt = statistics.median(times)
        exact = all(np.array_equal(a, b) for a, b in zip(w, (keys, sums, counts)))
        log(f"cpu baseline: {threads} threads, median {dt*1e3:.1f} ms, groups {w[0].shape[0]}, "
            f"equal to the device groups {exact}")
        out["cpu_baseline"] = {"value": args.rows / dt, "unit": "rows/s", "cores": threads, "kind": "port",
                               "sample": f"full workload ({args.rows} rows, decoded columns from the same seed), "
                                         f"star plan of 02inmem.hpp, median of {len(times)} runs",
                               "bit_exact_vs_gpu": bool(exact)}
        if not exact:
            raise SystemExit("star: device groups differ from the CPU plan's")
    print(json.dumps(out), flush=True)
    ctx.close()


def calibrate(ctx, args):
    """Three device copies of --calibrate-copy bytes in the profiled process: rocprofv3's
    FETCH_SIZE / WRITE_SIZE of that copy give the factors that turn the counters into bytes."""
    if args.calibrate_copy:
        a_, b_ = ctx.alloc(args.calibrate_copy), ctx.alloc(args.calibrate_copy)
        ctx.memset(a_, 1, args.calibrate_copy)
        for _ in range(3):
            ctx.copy(b_, a_, args.calibrate_copy)
        ctx.free(a_)
        ctx.free(b_)
        log(f"calibration: 3 copies of {args.calibrate_copy} bytes")


def traffic_of(name):
    """HBM bytes per launch of a workload's dominant kernel from its committed rocprof summary."""
    f = os.path.join(ROOT, "profiles", name + "_traffic.json")
    if os.path.exists(f):
        return json.load(open(f))["hbm_bytes_per_launch"]
    return None


def star_setup(args, chunk_rows):
    """The star schema of the star workloads: 600M fact rows (date_key sorted, bit-packed delta;
    customer, supplier, part keys uniform and revenue, LZ4), dimensions of 2,556 / 2,000,000 /
    200,000 / 1,400,000 rows with region (5), nation (25) and brand (1,000) attributes."""
    import golap_public_amd as g
    S = dict(n_cust=2_000_000, n_supp=200_000, n_part=1_400_000)
    key, year, _ = g.gen_date_dim(2556)
    log(f"generating {args.rows} star rows")
    fact = g.gen_star_fact(args.seed, args.rows, args.rows, key, S["n_cust"], S["n_supp"], S["n_part"])
    S["nation"] = g.gen_uniform_i32(args.seed, 20, S["n_cust"], 0, 24)
    S["region"] = g.gen_uniform_i32(args.seed, 21, S["n_supp"], 0, 4)
    S["brand"] = g.gen_uniform_i32(args.seed, 22, S["n_part"], 1, 1000)
    log("encoding")
    enc = dict(page_rows=args.page_rows, chunk_rows=chunk_rows, lz4_min_match=args.min_match)
    hosts = [g.HostColumn.encode(fact[0], codec=g.CODEC_DBP, **enc)] + \
            [g.HostColumn.encode(c, codec=g.CODEC_LZ4, **enc) for c in fact[1:]]
    del fact
    ctx = g.Context(0)
    S["cols"] = [ctx.upload(h) for h in hosts]
    S["comp"] = sum(h.desc().payload_bytes for h in hosts)
    del hosts
    dev = lambda a: ctx.to_device(np.ascontiguousarray(a))
    S.update(key=key, year=year, d_key=dev(key), d_year=dev(year.astype(np.int32)),
             c_key=dev(np.arange(S["n_cust"], dtype=np.int32)), c_nat=dev(S["nation"]),
             s_key=dev(np.arange(S["n_supp"], dtype=np.int32)), s_reg=dev(S["region"]),
             p_key=dev(np.arange(S["n_part"], dtype=np.int32)), p_brand=dev(S["brand"]))
    log(f"compressed payload {S['comp']/1e9:.3f} GB for {args.rows*20/1e9:.3f} GB raw")
    return ctx, S


def batch_queries(args, S, nch):
    """Seeded query mix: slice of log-uniform size in [1M, 600M] rows at a uniform start; date
    range covering 5..50 % of the 2,556 days (the sorted fact then prunes 50..95 % of the zones);
    supplier region, customer nation and a brand range of 1..10 brands drawn uniformly."""
    rng = np.random.default_rng(args.seed)
    qs = []
    zr = args.zone_rows
    for _ in range(args.queries):
        rows = float(np.exp(rng.uniform(np.log(1e6), np.log(min(6e8, args.rows)))))
        nc = max(1, min(nch, int(round(rows / zr))))
        cb = int(rng.integers(0, nch - nc + 1))
        width = int(2556 * rng.uniform(0.05, 0.5))
        d0 = int(rng.integers(0, 2556 - width + 1))
        bw = int(rng.integers(1, 11))
        b0 = int(rng.integers(1, 1000 - bw + 2))
        qs.append(dict(cb=cb, ce=cb + nc, dlo=int(S["key"][d0]), dhi=int(S["key"][d0 + width - 1]),
                       region=int(rng.integers(0, 5)), nation=int(rng.integers(0, 25)), blo=b0, bhi=b0 + bw - 1))
    return qs


def run_batch(args):
    """64 star queries over the date-sorted 600M-row star schema with min/max zone maps per
    65,536-row chunk: one persistent fused launch under a hipGraph (autotuned configuration)
    against one-query-at-a-time plain launches (zone-map check + golap_star_run per query),
    dimension maps resident for both. Every query's groups are compared between the two."""
    import golap_public_amd as g
    ctx, S = star_setup(args, args.zone_rows)
    cols = S["cols"]
    nch = ctx.column_desc(cols[0]).n_chunks
    qs = batch_queries(args, S, nch)
    log(f"{len(qs)} queries, {nch} zones; building their dimension maps")
    maps = []
    t0 = time.perf_counter()
    for q in qs:
        m = [ctx.hashmap(2556, S["d_key"], g.I32), ctx.hashmap(S["n_cust"], S["c_key"], g.I32),
             ctx.hashmap(S["n_supp"], S["s_key"], g.I32), ctx.hashmap(S["n_part"], S["p_key"], g.I32)]
        ctx.hashmap_build(m[0], 2556, [(S["d_key"], g.I32, q["dlo"], q["dhi"])], S["d_year"])
        ctx.hashmap_build(m[1], S["n_cust"], [(S["c_nat"], g.I32, q["nation"], q["nation"])])
        ctx.hashmap_build(m[2], S["n_supp"], [(S["s_reg"], g.I32, q["region"], q["region"])])
        ctx.hashmap_build(m[3], S["n_part"], [(S["p_brand"], g.I32, q["blo"], q["bhi"])], S["p_brand"])
        maps.append(m)
    ctx.sync()
    build_ms = (time.perf_counter() - t0) * 1e3
    dims_of = lambda i: [(cols[0], maps[i][0], 1992, 7), (cols[1], maps[i][1], 0, 0),
                         (cols[2], maps[i][2], 0, 0), (cols[3], maps[i][3], 1, 1000)]
    flags = ctx.alloc(nch * 2)

    def plain_all():
        out = []
        for i, q in enumerate(qs):
            ctx.minmax_check(cols[0], q["dlo"], q["dhi"], flags)
            out.append(ctx.star_run(dims_of(i), cols[4], chunk_begin=q["cb"], chunk_end=q["ce"],
                                    dev_chunk_mask=flags))
        return out

    entries = [dict(dims=dims_of(i), value=cols[4], chunk_begin=q["cb"], chunk_end=q["ce"],
                    zone=(cols[0], q["dlo"], q["dhi"])) for i, q in enumerate(qs)]
    b = ctx.star_batch(entries)
    default_cfg = b.config().as_dict()
    info = b.run()
    t0 = time.perf_counter()
    best, tuned_ms = b.tune(reps=3)
    tune_s = time.perf_counter() - t0
    log(f"autotuner: {default_cfg} -> {best.as_dict()} ({tuned_ms:.3f} ms, {tune_s:.1f} s)")
    # the shared-filter schedule under the tuned grain and occupancy, for the record
    shared_ms = None
    try:
        b.set_config(**dict(best.as_dict(), shared_filter=1))
        shared_ms = statistics.median(b.run().ms for _ in range(3))
    except g.GolapError:
        pass
    b.set_config(**best.as_dict())
    log(f"shared-filter schedule: {shared_ms} ms")
    for _ in range(args.warmup):
        b.run()
        plain_all()
    # batched: graph replay + every query's groups back on the host
    dev_ms, walls, run_walls = [], [], []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        info = b.run(use_graph=True)
        t2 = time.perf_counter()
        res = b.results()
        walls.append(time.perf_counter() - t1)
        run_walls.append(t2 - t1)
        dev_ms.append(info.ms)
    plain_walls, plain_dev = [], []
    for _ in range(max(1, args.steps // 4)):
        t1 = time.perf_counter()
        ref = plain_all()
        plain_walls.append(time.perf_counter() - t1)
        plain_dev.append(sum(r[0].kernel_ms for r in ref))
    for i in range(len(qs)):
        assert all(np.array_equal(x, y) for x, y in zip(res[i][1:], ref[i][1:])), f"query {i} differs"
    wall = statistics.median(walls)
    pwall = statistics.median(plain_walls)
    dms = statistics.median(dev_ms)
    surv_rows = info.pages * args.page_rows
    log(f"batch: {dms:.3f} ms device, {wall*1e3:.3f} ms with results; plain: {pwall*1e3:.3f} ms "
        f"({statistics.median(plain_dev):.3f} ms kernels); items {info.items}, surviving pages {info.pages}")
    gbps = info.comp_bytes / (dms * 1e-3) / 1e9
    out = {"metric": "queries/s (64-query batch, one persistent fused launch, 1 GPU)",
           "value": len(qs) / wall, "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "int64", "data": "synthetic: SplitMix64-seeded star schema, seeded queries",
           "config": {"workload": "star_batch_64q_600M", "rows": args.rows, "zone_rows": args.zone_rows,
                      "page_rows": args.page_rows, "queries": len(qs), "zones": nch,
                      "surviving_pages": info.pages, "surviving_rows": surv_rows, "work_items": info.items,
                      "comp_bytes": info.comp_bytes, "device_ms": dms,
                      "fact_rows_per_s": surv_rows / (dms * 1e-3),
                      "plain_queries_per_s": len(qs) / pwall, "plain_ms": pwall * 1e3,
                      "plain_kernel_ms": statistics.median(plain_dev),
                      "speedup_vs_plain": pwall / wall, "map_build_ms_all_queries": build_ms,
                      "run_ms": statistics.median(run_walls) * 1e3,
                      "results_ms": (wall - statistics.median(run_walls)) * 1e3,
                      "tuned_config": best.as_dict(), "default_config": default_cfg, "tune_s": tune_s,
                      "shared_filter_ms": shared_ms},
           "roofline": {"bound": "hbm", "achieved": gbps, "peak": 8000.0, "unit": "GB/s", "frac": gbps / 8000.0,
                        "traffic": traffic_of("star_batch_64q_600M")}}
    calibrate(ctx, args)
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = len(os.sched_getaffinity(0))
        sample = min(args.rows, 60_000_000)
        f2 = oracle.gen_star_fact(args.seed, sample, args.rows, S["key"], S["n_cust"], S["n_supp"], S["n_part"])
        q = qs[0]
        t0 = time.perf_counter()
        oracle.star_groupby([(S["key"], [(S["key"], q["dlo"], q["dhi"])], S["year"].astype(np.int32)),
                             (np.arange(S["n_cust"], dtype=np.int32), [(S["nation"], q["nation"], q["nation"])], None),
                             (np.arange(S["n_supp"], dtype=np.int32), [(S["region"], q["region"], q["region"])], None),
                             (np.arange(S["n_part"], dtype=np.int32), [(S["brand"], q["blo"], q["bhi"])], S["brand"])],
                            list(f2[:4]), f2[4], threads=threads)
        dt = time.perf_counter() - t0
        rows_per_q = surv_rows / len(qs)
        out["cpu_baseline"] = {"value": sample / dt / rows_per_q, "unit": "queries/s", "cores": threads,
                               "kind": "port",
                               "sample": f"star plan of 02inmem.hpp over the first {sample} decoded rows "
                                         f"({sample/dt:.3g} rows/s), scaled to this mix's mean "
                                         f"{rows_per_q:.3g} surviving rows per query"}
    print(json.dumps(out), flush=True)
    b.free()
    ctx.close()


TRIP_HOURS = 8784


def trip_template(t, h0, span, t0):
    """Request template t over hours [h0, h0 + span) of the year: 0 SUM(fare), 1 passengers >= 2
    SUM(distance), 2 fare in [1000, 5000] SUM(distance), 3 distance >= 5000 COUNT by
    (passenger_count, zone); the others group by zone x hour of the range (up to 265 x 8,784 =
    2,327,760 groups). Columns: 0 ts, 1 passenger, 2 zone, 3 fare, 4 distance."""
    lo, hi = t0 + h0 * 3600, t0 + (h0 + span) * 3600 - 1
    zone, hour = (2, 1, 1, 265, 1), (0, t0 + h0 * 3600, 3600, span, h0)
    if t == 0:
        return dict(preds=[(0, lo, hi)], keys=[zone, hour], value_col=3)
    if t == 1:
        return dict(preds=[(0, lo, hi), (1, 2, 6)], keys=[zone, hour], value_col=4)
    if t == 2:
        return dict(preds=[(0, lo, hi), (3, 1000, 5000)], keys=[hour, zone], value_col=4)
    return dict(preds=[(0, lo, hi), (4, 5000, 1 << 30)], keys=[(1, 1, 1, 6, 1), zone], value_col=-1)


def trip_request(rng, t0):
    """A seeded request: a uniform template over a time range of log-uniform length
    (1 hour .. 1 year)."""
    t = int(rng.integers(0, 4))
    span = int(np.exp(rng.uniform(0, np.log(TRIP_HOURS))))
    h0 = int(rng.integers(0, TRIP_HOURS - span + 1))
    return trip_template(t, h0, span, t0)


class ServeRig:
    """The serving workload: the trip table resident in HBM (generated on the host, coded on the
    device by golap_encode_device, byte-identical to the host encoder), seeded requests drawn
    from the four templates, and the two ways of running them: the request loop (HBM-budget
    admission, continuous batching into persistent fused launches on two streams) and
    one-query-at-a-time plain launches (golap_scan_run). Latency = arrival to results on the
    host."""

    OPTS = dict(slots=2, max_batch=64, emit_threads=8)

    def __init__(self, args, ctx, n_requests, keep_cpu_rows=0):
        import golap_public_amd as g
        self.g, self.ctx, self.args = g, ctx, args
        n = args.trip_rows
        t_setup = time.perf_counter()
        log(f"serve: generating {n} trip rows")
        cols = g.gen_trip(args.seed, n)
        enc = dict(page_rows=args.page_rows, chunk_rows=args.zone_rows, lz4_min_match=args.min_match)
        codecs = [g.CODEC_DBP, g.CODEC_LZ4, g.CODEC_FOR, g.CODEC_FOR, g.CODEC_FOR]
        self.cpu_cols = [c[:keep_cpu_rows].copy() for c in cols] if keep_cpu_rows else None
        self.dev, self.comp = [], 0
        for c, k in zip(cols, codecs):
            dv = ctx.to_device(c)
            col, _ = ctx.encode_device(dv, c.size, g.TYPE_OF[c.dtype], codec=k, **enc)
            ctx.free(dv)
            self.dev.append(col)
            self.comp += ctx.column_desc(col).payload_bytes
        del cols
        self.n = n
        log(f"serve: compressed {self.comp/1e9:.3f} GB for {n*19/1e9:.3f} GB raw; setup "
            f"{time.perf_counter()-t_setup:.1f} s")
        rng = np.random.default_rng(args.seed)
        self.reqs = [dict(cols=self.dev, **trip_request(rng, g.TRIP_T0)) for _ in range(n_requests)]
        self.plans = [ctx.scan_plan(**r) for r in self.reqs]
        self.doms = [int(np.prod([k[3] for k in r["keys"]])) for r in self.reqs]
        self.maxd = max(self.doms)
        self.pb = self.bufs()
        for i in range(min(8, n_requests)):  # warm-up
            self.plain(i)

    def bufs(self):
        return (np.empty(self.maxd * 2, np.int32), np.empty(self.maxd, np.int64), np.empty(self.maxd, np.uint64))

    def plain(self, i):
        import ctypes as C
        g, ctx = self.g, self.ctx
        r = g.StarResult()
        ctx.check(ctx.lib.golap_scan_run(ctx.h, C.byref(self.plans[i]), C.byref(r), g._ptr(self.pb[0]),
                                         g._ptr(self.pb[1]), g._ptr(self.pb[2]), self.doms[i]), "scan_run")
        return r

    def plain_capacity(self, k=200):
        """One-at-a-time plain launches back to back (queries/s)."""
        k = min(k, len(self.reqs))
        t0 = time.perf_counter()
        for i in range(k):
            self.plain(i)
        return k / (time.perf_counter() - t0)

    def served(self, arrivals):
        """Requests 0..len(arrivals)-1 submitted at their arrival times (seconds from the start;
        all at 0 for a closed burst). Returns (latencies ms, results, wall s, server stats)."""
        import ctypes as C
        import threading
        g, ctx = self.g, self.ctx
        m = len(arrivals)
        lat, res, chunks = [0.0] * m, [None] * m, [0] * m
        tickets, sub_delay = [None] * m, [0.0] * m
        with ctx.server(hbm_budget=self.args.budget_mb << 20, **self.OPTS) as srv:
            def waiter(w, nw):
                b = self.bufs()
                for i in range(w, m, nw):
                    while tickets[i] is None:
                        time.sleep(0.0001)
                    r, lt = g.StarResult(), C.c_double()
                    ctx.check(ctx.lib.golap_server_wait(srv.h, tickets[i], C.byref(r), g._ptr(b[0]), g._ptr(b[1]),
                                                        g._ptr(b[2]), self.doms[i], C.byref(lt)), "server_wait")
                    lat[i] = lt.value + sub_delay[i]
                    res[i] = (r.n_groups, r.total_sum, r.rows_matched)
                    chunks[i] = r.chunks_run

            nw = 8
            ws = [threading.Thread(target=waiter, args=(w, nw)) for w in range(nw)]
            for w in ws:
                w.start()
            start = time.perf_counter()
            for i in range(m):
                now = time.perf_counter() - start
                if arrivals[i] > now:
                    time.sleep(arrivals[i] - now)
                t = C.c_uint64()
                ctx.check(ctx.lib.golap_server_submit_scan(srv.h, C.byref(self.plans[i]), C.byref(t)), "server_submit")
                sub_delay[i] = max(0.0, (time.perf_counter() - start) - arrivals[i]) * 1e3  # late submits count
                tickets[i] = t.value
            for w in ws:
                w.join()
            wall = time.perf_counter() - start
            st = srv.stats()
        self.chunks = chunks
        return lat, res, wall, st

    def plain_at(self, arrivals, want):
        """Plain launches at the same arrivals (one worker, arrival order), each result checked
        against the served one. Returns (latencies ms, wall s)."""
        lat = [0.0] * len(arrivals)
        start = time.perf_counter()
        for i in range(len(arrivals)):
            now = time.perf_counter() - start
            if arrivals[i] > now:
                time.sleep(arrivals[i] - now)
            r = self.plain(i)
            lat[i] = (time.perf_counter() - start - arrivals[i]) * 1e3
            assert (r.n_groups, r.total_sum, r.rows_matched) == want[i], f"request {i} differs"
        return lat, time.perf_counter() - start

    def sweep(self, n_req):
        """Poisson arrivals of n_req requests at 0.5x and 0.9x the plain capacity, at 0.9x the
        served capacity (a closed burst of n_req requests) and at 3.2x the plain capacity, both
        forms at each rate. A form keeps up with a rate when the median latency of the last
        quarter of the requests is less than twice that of the first quarter (a growing queue
        shows as growth whatever the run length)."""
        n_req = min(n_req, len(self.reqs))
        plain_cap = self.plain_capacity()
        _, _, burst_wall, _ = self.served([0.0] * n_req)
        served_cap = n_req / burst_wall
        log(f"serve: plain capacity {plain_cap:.1f} q/s, served capacity {served_cap:.1f} q/s")
        pct = lambda a, q: float(np.percentile(np.array(a), q))
        rows, last_st = [], None
        q = max(1, n_req // 4)
        growth = lambda a: float(np.median(np.array(a[-q:])) / max(np.median(np.array(a[:q])), 1e-3))
        for name, rate in (("0.5x plain capacity", 0.5 * plain_cap), ("0.9x plain capacity", 0.9 * plain_cap),
                           ("0.9x served capacity", 0.9 * served_cap), ("3.2x plain capacity", 3.2 * plain_cap)):
            arr = np.cumsum(np.random.default_rng(self.args.seed + 1).exponential(1.0 / rate, n_req))
            ls, res, ws, st = self.served(arr)
            lp, wp = self.plain_at(arr, res)
            row = {"rate": name, "offered_qps": rate, "requests": n_req,
                   "served_qps": n_req / ws, "served_p50_ms": pct(ls, 50), "served_p99_ms": pct(ls, 99),
                   "plain_qps": n_req / wp, "plain_p50_ms": pct(lp, 50), "plain_p99_ms": pct(lp, 99)}
            row["served_over_plain_qps"] = row["served_qps"] / row["plain_qps"]
            row["served_latency_growth"], row["plain_latency_growth"] = growth(ls), growth(lp)
            row["served_keeps_up"] = row["served_latency_growth"] < 2.0
            row["plain_keeps_up"] = row["plain_latency_growth"] < 2.0
            row["meets_3x_and_lower_p99"] = bool(row["served_over_plain_qps"] >= 3.0 and
                                                 row["served_p99_ms"] < row["plain_p99_ms"])
            log(f"serve @ {name} ({rate:.1f} q/s): served {row['served_qps']:.1f} q/s p50 "
                f"{row['served_p50_ms']:.2f} p99 {row['served_p99_ms']:.2f} ms; plain {row['plain_qps']:.1f} q/s "
                f"p50 {row['plain_p50_ms']:.2f} p99 {row['plain_p99_ms']:.2f} ms; latency growth served "
                f"{row['served_latency_growth']:.2f} plain {row['plain_latency_growth']:.2f}")
            rows.append(row)
            last_st = st
        return {"plain_capacity_qps": plain_cap, "served_capacity_qps": served_cap, "rates": rows}, last_st

    def free(self):
        for d in self.dev:
            self.ctx.free_column(d)


def run_serve(args):
    """Serving form on the trip table: the sweep of ServeRig.sweep; the line's value is the served
    queries/s at 3.2x the plain capacity (a rate plain launches cannot keep up with)."""
    import golap_public_amd as g
    ctx = g.Context(0)
    cpu_rows = 0 if args.no_cpu_baseline else min(args.trip_rows, 150_000_000)
    rig = ServeRig(args, ctx, args.requests, keep_cpu_rows=cpu_rows)
    sw, st = rig.sweep(args.requests)
    top = sw["rates"][-1]
    # algorithmic bytes: every admitted request's non-pruned compressed bytes; the shared launches
    # read each surviving chunk's columns once for all requests of a batch (shared_comp_bytes)
    gbps = st["comp_bytes"] / (st["device_ms"] * 1e-3) / 1e9 if st["device_ms"] else 0.0
    out = {"metric": "queries/s + p99 latency (served trip-table requests, Poisson arrivals, 1 GPU)",
           "value": top["served_qps"], "unit": "queries/s", "n_gpus": 1, "steps": top["requests"], "warmup": 8,
           "ms_per_step": 1e3 / top["served_qps"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "int64", "data": "synthetic: SplitMix64-seeded trip table, seeded requests",
           "config": {"workload": "serve_trip_1.5B", "rows": rig.n, "zone_rows": args.zone_rows,
                      "page_rows": args.page_rows, "lz4_min_match": args.min_match,
                      "codecs": "ts DBP, passenger LZ4, zone/fare/distance FOR", "comp_bytes_table": rig.comp,
                      "sweep": sw, "server": st, "server_opts": dict(hbm_budget=args.budget_mb << 20, **rig.OPTS)},
           "roofline": {"bound": "hbm", "achieved": gbps, "peak": 8000.0, "unit": "GB/s", "frac": gbps / 8000.0,
                        "traffic": traffic_of("serve_trip_1.5B")}}
    calibrate(ctx, args)
    if cpu_rows:
        # the four templates over the whole year on the first cpu_rows rows (every row passes the
        # time filter), median of 3, scaled to the mix's mean scanned rows per request
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = len(os.sched_getaffinity(0))
        per_t = []
        for t in range(4):
            q = trip_template(t, 0, TRIP_HOURS, g.TRIP_T0)
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                w = oracle.scan_groupby(rig.cpu_cols, q["preds"], q["keys"], q["value_col"], threads=threads)
                ts.append(time.perf_counter() - t0)
                assert w is not None
            per_t.append(statistics.median(ts))
        rows_s = 4 * cpu_rows / sum(per_t)
        rows_per_req = statistics.mean(rig.chunks) * args.zone_rows
        out["cpu_baseline"] = {"value": rows_s / rows_per_req, "unit": "queries/s", "cores": threads,
                               "kind": "port",
                               "sample": f"the four request templates over the first {cpu_rows} decoded rows "
                                         f"({rows_s:.3g} rows/s, median of 3), scaled to the last rate's mean "
                                         f"{rows_per_req:.3g} scanned rows per request"}
    print(json.dumps(out), flush=True)
    rig.free()
    ctx.close()


def run_q1_files(args):
    """The q1 plan streamed from four column files (windows of chunks read, copied to HBM and
    scanned in a pipeline); each step is a whole query from open to result. The files were just
    written, so reads mostly come from the page cache: the rate is that of the host read path,
    not of a cold SSD."""
    import shutil
    import tempfile

    import golap_public_amd as g

    key, year, _ = g.gen_date_dim(2556)
    log(f"generating {args.rows} rows")
    cols = g.gen_q1_fact(args.seed, args.rows, key)
    hosts = [g.HostColumn.encode(c, codec=g.CODEC_LZ4, page_rows=args.page_rows, chunk_rows=args.chunk_rows,
                                 lz4_min_match=args.min_match) for c in cols]
    del cols
    folder = args.stream_dir or tempfile.mkdtemp(prefix="golap_q1_files_")
    os.makedirs(folder, exist_ok=True)
    paths = [os.path.join(folder, f"{n}.gcol") for n in ("date", "disc", "qty", "price")]
    ctx = g.Context(0)
    try:
        t = time.perf_counter()
        for h, p in zip(hosts, paths):
            h.save(p)
        file_bytes = sum(os.path.getsize(p) for p in paths)
        log(f"wrote {file_bytes/1e9:.3f} GB in {time.perf_counter()-t:.1f} s to {folder}")
        cus, name = ctx.info()
        dk, dy = ctx.to_device(key), ctx.to_device(year)
        m = ctx.hashmap(key.size, dk, g.I32)
        ctx.hashmap_build(m, key.size, [(dy, g.I16, 1995, 1995)])
        dcols = [ctx.upload(h) for h in hosts]
        want = ctx.q1_run(*dcols, m, (4, 6), (30, 50))
        for c in dcols:
            ctx.free_column(c)
        del hosts
        for _ in range(args.warmup):
            r, st = ctx.q1_run_files(paths, m, (4, 6), (30, 50))
        steps = []
        for i in range(args.steps):
            r, st = ctx.q1_run_files(paths, m, (4, 6), (30, 50))
            if (r.sum, r.matches) != (want.sum, want.matches):
                raise SystemExit(f"q1_files: streamed result {r.sum} differs from the resident one {want.sum}")
            steps.append((st.total_ms, st.read_ms, r.kernel_ms))
            log(f"step {i}: {st.total_ms:.1f} ms, reads {st.read_ms:.1f} ms, kernels {r.kernel_ms:.2f} ms, "
                f"{st.windows} windows, direct_io {st.direct_io}")
        ms = statistics.median(x[0] for x in steps)
        out = {
            "metric": "q1 rows per second streamed from column files", "value": args.rows / (ms * 1e-3),
            "unit": "rows/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
            "data": "synthetic: the q1 workload's columns, LZ4 pages, written to column files",
            "config": {"workload": "q1_files_600M_lz4", "rows": args.rows, "chunk_rows": args.chunk_rows,
                       "file_bytes": file_bytes, "bytes_read": st.bytes_read, "windows": st.windows,
                       "window_chunks": st.window_chunks, "direct_io": st.direct_io,
                       "read_ms": statistics.median(x[1] for x in steps),
                       "kernel_ms": statistics.median(x[2] for x in steps),
                       "read_GBps": st.bytes_read / (ms * 1e-3) / 1e9,
                       "result_sum": r.sum, "result_matches": r.matches, "device": name},
        }
        print(json.dumps(out), flush=True)
    finally:
        ctx.close()
        if not args.stream_dir:
            shutil.rmtree(folder, ignore_errors=True)


def run_encode(args):
    """The write path on the device: the q1 workload's four columns (600M rows, 6 GB of values
    resident in HBM) coded into LZ4 pages with zone maps by golap_encode_device; one step codes
    all four. The device columns are compared byte for byte with the host encoder's once. The
    CPU baseline is the host encoder (all host threads) on the same values."""
    import golap_public_amd as g

    key, _, _ = g.gen_date_dim(2556)
    log(f"generating {args.rows} rows")
    cols = g.gen_q1_fact(args.seed, args.rows, key)
    ctx = g.Context(0)
    cus, name = ctx.info()
    opts = dict(codec=g.CODEC_LZ4, page_rows=args.page_rows, chunk_rows=args.chunk_rows,
                lz4_min_match=args.min_match)
    raw = sum(c.nbytes for c in cols)
    dv = [ctx.to_device(c) for c in cols]

    def step():
        out, ms = [], 0.0
        for c, p in zip(cols, dv):
            col, k = ctx.encode_device(p, c.size, g.TYPE_OF[c.dtype], **opts)
            out.append(col)
            ms += k
        return out, ms

    def drop(out):
        for col in out:
            ctx.free_column(col)

    for _ in range(args.warmup):
        drop(step()[0])
    kms = []
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        out, ms = step()
        kms.append(ms)
        if i < args.steps - 1:
            drop(out)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    kernel_ms = statistics.mean(kms)
    comp = sum(ctx.column_desc(c).comp_bytes for c in out)
    # byte-for-byte against the host encoder, which is also the CPU baseline
    t = time.perf_counter()
    hosts = [g.HostColumn.encode(c, **opts) for c in cols]
    host_s = time.perf_counter() - t
    same = all(np.array_equal(ctx.to_host(ctx.column_desc(c).comp, ctx.column_desc(c).comp_bytes, np.uint8),
                              h.arrays()["comp"]) for c, h in zip(out, hosts))
    drop(out)
    if not same:
        raise SystemExit("encode: device columns differ from the host encoder's")
    log(f"device {kernel_ms:.1f} ms kernels per step ({raw / (kernel_ms * 1e-3) / 1e9:.1f} GB/s of values), "
        f"{elapsed * 1e3 / args.steps:.1f} ms per step; host encoder {host_s * 1e3:.0f} ms; {comp / 1e9:.3f} GB coded")
    ms_per_step = elapsed * 1e3 / args.steps
    out = {
        "metric": "values coded per second (LZ4 pages + zone maps)", "value": raw / (ms_per_step * 1e-3),
        "unit": "B/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic: the q1 workload's columns resident in HBM",
        "config": {"workload": "encode_q1_600M_lz4", "rows": args.rows, "raw_bytes": raw, "comp_bytes": comp,
                   "kernel_ms": kernel_ms, "byte_identical_to_host": same, "device": name},
        "cpu_baseline": {"value": raw / host_s, "unit": "B/s", "cores": len(os.sched_getaffinity(0)),
                         "kind": "port", "sample": "the host encoder on the same four columns, one run"},
    }
    print(json.dumps(out), flush=True)
    for p in dv:
        ctx.free(p)
    ctx.close()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this library runs on one GPU; --gpus {args.gpus} is not supported "
              "(each GPU would scan its own share of the chunks)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "star":
        return run_star(args)
    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "serve":
        return run_serve(args)
    if args.workload == "q1_files":
        return run_q1_files(args)
    if args.workload == "encode":
        return run_encode(args)
    import golap_public_amd as g

    t_setup = time.perf_counter()
    key, year, _ = g.gen_date_dim(2556)
    log(f"generating {args.rows} rows")
    cols = g.gen_q1_fact(args.seed, args.rows, key)
    log("encoding")
    enc = dict(codec=g.CODEC_LZ4, page_rows=args.page_rows, chunk_rows=args.chunk_rows)
    hosts = [g.HostColumn.encode(c, lz4_min_match=args.min_match, lz4_unit=args.lz4_unit, **enc) for c in cols]
    raw_bytes = args.rows * 10
    comp_payload = sum(h.desc().payload_bytes for h in hosts)
    log(f"compressed payload {comp_payload/1e9:.3f} GB for {raw_bytes/1e9:.3f} GB raw")
    ctx = g.Context(0)
    cus, name = ctx.info()
    dcols = [ctx.upload(h) for h in hosts]
    # the same columns coded with minimum match 16 (on the device; byte-identical to the host
    # encoder): a secondary line where most pages are one literal run, read in place
    mm16 = []
    if not args.no_mm16 and args.min_match != 16:
        for c in cols:
            dv = ctx.to_device(c)
            mm16.append(ctx.encode_device(dv, c.size, g.TYPE_OF[c.dtype], lz4_min_match=16, **enc)[0])
            ctx.free(dv)
    del cols
    dk, dy = ctx.to_device(key), ctx.to_device(year)
    m = ctx.hashmap(key.size, dk, g.I32)
    log(f"setup {time.perf_counter()-t_setup:.1f} s on {name} ({cus} CUs)")

    def step(cols=dcols):
        ctx.hashmap_build(m, key.size, [(dy, g.I16, 1995, 1995)])
        return ctx.q1_run(*cols, m, (4, 6), (30, 50))

    for _ in range(args.warmup):
        r = step()
    ctx.sync()
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = step()
        kms.append(r.kernel_ms)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = args.rows * args.steps / elapsed
    kernel_ms = statistics.mean(kms)
    algo_bytes = r.comp_bytes + key.size * (4 + 2) + 24
    achieved = algo_bytes / (kernel_ms * 1e-3) / 1e9
    traffic = None
    if os.path.exists(args.traffic_file):
        try:
            with open(args.traffic_file) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    log(f"sum {r.sum} matches {r.matches}; {ms_per_step:.3f} ms/step, kernel {kernel_ms:.3f} ms, "
        f"{achieved:.0f} GB/s")
    # what a plain streaming read of 4 GiB reaches on this device (best of 5), beside the peak
    probe_bytes = 4 << 30
    pbuf = ctx.alloc(probe_bytes)
    ctx.memset(pbuf, 1, probe_bytes)
    probe_ms = min(ctx.read_probe(pbuf, probe_bytes) for _ in range(6))
    ctx.free(pbuf)
    probe_gbps = probe_bytes / (probe_ms * 1e-3) / 1e9
    log(f"plain streaming read: {probe_gbps:.0f} GB/s; the fused kernel's HBM traffic moves at "
        f"{(traffic or algo_bytes) / (kernel_ms * 1e-3) / 1e9:.0f} GB/s")

    out = {
        "metric": METRIC, "value": value, "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int64",
        "data": "synthetic: SplitMix64-seeded columns with the workload's distributions, LZ4 pages",
        "config": {"workload": "q1_fused_600M_lz4", "rows": args.rows, "chunk_rows": args.chunk_rows,
                   "page_rows": args.page_rows, "codec": "lz4", "lz4_min_match": args.min_match,
                   "lz4_unit": "bytes" if args.lz4_unit == 1 else "element", "comp_bytes": r.comp_bytes, "raw_bytes": raw_bytes,
                   "plan": "year=1995, discount 4..6, quantity>=30, SUM(price*discount)",
                   "result_sum": r.sum, "result_matches": r.matches, "kernel_ms": kernel_ms,
                   "device": name, "read_probe_GBps": probe_gbps,
                   "hbm_traffic_GBps": (traffic / (kernel_ms * 1e-3) / 1e9) if traffic else None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
    }
    if mm16:
        for _ in range(args.warmup):
            r16 = step(mm16)
        k16 = []
        for _ in range(args.steps):
            r16 = step(mm16)
            k16.append(r16.kernel_ms)
        ms16 = statistics.mean(k16)
        a16 = (r16.comp_bytes + key.size * (4 + 2) + 24) / (ms16 * 1e-3) / 1e9
        assert (r16.sum, r16.matches) == (r.sum, r.matches)
        out["config"]["lz4_min_match_16"] = {
            "comp_bytes": r16.comp_bytes, "kernel_ms": ms16, "achieved_GBps": a16, "frac": a16 / HBM_PEAK_GBS,
            "rows_per_s_kernel": args.rows / (ms16 * 1e-3), "same_result": True}
        log(f"min-match 16 pages: kernel {ms16:.3f} ms, {a16:.0f} GB/s")
        for c in mm16:
            ctx.free_column(c)
    calibrate(ctx, args)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args, hosts, key, year, r.sum)
    del hosts
    if not args.no_serve_sweep:
        rig = ServeRig(args, ctx, args.sweep_requests)
        sw, _ = rig.sweep(args.sweep_requests)
        top = sw["rates"][-1]
        out["config"]["serving"] = {
            "workload": f"serve_trip_1.5B sweep ({rig.n} trip rows, {args.sweep_requests} requests per rate)",
            "served_over_plain_qps": top["served_over_plain_qps"], "served_p99_ms": top["served_p99_ms"],
            "plain_p99_ms": top["plain_p99_ms"], "at_rate": top["rate"], "served_keeps_up": top["served_keeps_up"],
            "plain_keeps_up": top["plain_keeps_up"], "sweep": sw}
        rig.free()
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
