"""Benchmark: the fused compressed-chunk -> filter -> join-probe -> int64-sum pipeline on one GPU.

Workload (one "step" = one query): 600,000,000 fact rows (date_key int32 over 2,556 days,
discount uint8 U[0,10], quantity uint8 U[1,50], price_cents int32 U[90000,10500000]) stored as
LZ4 pages in 262,144-row chunks, resident in HBM. Each step builds the date hash table
(year == 1995) on the device and runs the fused kernel for
discount in [4,6] and quantity >= 30, SUM(price*discount) as int64.

Prints one JSON line (see README of the driver contract): rows/s over the whole job, the
roofline of the fused kernel (algorithmic bytes / kernel time, against 8 TB/s), and the CPU
path (oracle decode + plan on the host cores) timed on the same data.
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
    ap.add_argument("--workload", choices=["q1", "star"], default="q1",
                    help="q1: fused filter/probe/sum (the headline line); star: star join + group-by")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=600_000_000)
    ap.add_argument("--page-rows", type=int, default=4096)
    ap.add_argument("--chunk-rows", type=int, default=262144)
    ap.add_argument("--min-match", type=int, default=16)
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
    return {"value": d0.n_rows / med, "unit": "rows/s", "cores": threads, "kind": "port",
            "sample": f"full workload ({d0.n_rows} rows, same compressed pages), median of "
                      f"{args.cpu_runs} runs, CPU: {model}",
            "bit_exact_vs_gpu": bool(res == gpu_sum)}


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
                      "comp_bytes_per_s_GBps": res.comp_bytes / (kernel_ms * 1e-3) / 1e9}}
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = len(os.sched_getaffinity(0))
        sample = min(args.rows, 60_000_000)
        f2 = oracle.gen_star_fact(args.seed, sample, args.rows, key, n_cust, n_supp, n_part)
        t0 = time.perf_counter()
        w = oracle.star_groupby([(key, [], year.astype(np.int32)),
                                 (np.arange(n_cust, dtype=np.int32), [(nation, 7, 7)], None),
                                 (np.arange(n_supp, dtype=np.int32), [(region, 1, 1)], None),
                                 (np.arange(n_part, dtype=np.int32), [(brand, 42, 42)], brand)],
                                list(f2[:4]), f2[4], threads=threads)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": sample / dt, "unit": "rows/s", "cores": threads, "kind": "port",
                               "sample": f"first {sample} rows (decoded columns), star plan of 02inmem.hpp"}
    print(json.dumps(out), flush=True)
    ctx.close()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the sharded form is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "star":
        return run_star(args)
    import golap_public_amd as g

    t_setup = time.perf_counter()
    key, year, _ = g.gen_date_dim(2556)
    log(f"generating {args.rows} rows")
    cols = g.gen_q1_fact(args.seed, args.rows, key)
    log("encoding")
    hosts = [g.HostColumn.encode(c, codec=g.CODEC_LZ4, page_rows=args.page_rows, chunk_rows=args.chunk_rows,
                                 lz4_min_match=args.min_match) for c in cols]
    del cols
    raw_bytes = args.rows * 10
    comp_payload = sum(h.desc().payload_bytes for h in hosts)
    log(f"compressed payload {comp_payload/1e9:.3f} GB for {raw_bytes/1e9:.3f} GB raw")
    ctx = g.Context(0)
    cus, name = ctx.info()
    dcols = [ctx.upload(h) for h in hosts]
    dk, dy = ctx.to_device(key), ctx.to_device(year)
    m = ctx.hashmap(key.size, dk, g.I32)
    log(f"setup {time.perf_counter()-t_setup:.1f} s on {name} ({cus} CUs)")

    def step():
        ctx.hashmap_build(m, key.size, [(dy, g.I16, 1995, 1995)])
        return ctx.q1_run(*dcols, m, (4, 6), (30, 50))

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

    out = {
        "metric": METRIC, "value": value, "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int64",
        "data": "synthetic: SplitMix64-seeded columns with the workload's distributions, LZ4 pages",
        "config": {"workload": "q1_fused_600M_lz4", "rows": args.rows, "chunk_rows": args.chunk_rows,
                   "page_rows": args.page_rows, "codec": "lz4", "lz4_min_match": args.min_match,
                   "comp_bytes": r.comp_bytes, "raw_bytes": raw_bytes,
                   "plan": "year=1995, discount 4..6, quantity>=30, SUM(price*discount)",
                   "result_sum": r.sum, "result_matches": r.matches, "kernel_ms": kernel_ms,
                   "device": name},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
    }
    if args.calibrate_copy:
        a_, b_ = ctx.alloc(args.calibrate_copy), ctx.alloc(args.calibrate_copy)
        ctx.memset(a_, 1, args.calibrate_copy)
        for _ in range(3):
            ctx.copy(b_, a_, args.calibrate_copy)
        ctx.free(a_)
        ctx.free(b_)
        log(f"calibration: 3 copies of {args.calibrate_copy} bytes")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args, hosts, key, year, r.sum)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
