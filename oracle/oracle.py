"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, always as the checker or the CPU baseline, never as the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

vp, u64, u32, i64 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int64


class SsbGolden(C.Structure):
    _fields_ = [("lineorder_rows", u64), ("q1", u64 * 3), ("q2_sum", u64 * 3), ("q2_groups", u64 * 3),
                ("q34_sum", u64 * 7), ("q34_groups", u64 * 7)]


SSB_COLS = [("lo_orderdate", "lo", "orderdate"), ("lo_discount", "lo", "discount"), ("lo_quantity", "lo", "quantity"),
            ("lo_extendedprice", "lo", "extendedprice"), ("lo_partkey", "lo", "partkey"),
            ("lo_suppkey", "lo", "suppkey"), ("lo_custkey", "lo", "custkey"), ("lo_revenue", "lo", "revenue"),
            ("lo_supplycost", "lo", "supplycost"), ("d_key", "d", "key"), ("d_year", "d", "year"),
            ("d_yearmonthnum", "d", "yearmonthnum"), ("d_weeknuminyear", "d", "weeknuminyear"),
            ("p_category", "p", "category"), ("p_brand1", "p", "brand1"), ("p_mfgr", "p", "mfgr"),
            ("s_region", "s", "region"), ("s_nation", "s", "nation"), ("s_city", "s", "city"),
            ("c_region", "c", "region"), ("c_nation", "c", "nation"), ("c_city", "c", "city")]


class SsbCols(C.Structure):
    _fields_ = [(name, vp) for name, _, _ in SSB_COLS]


class Pred(C.Structure):
    _fields_ = [("col", vp), ("type", u32), ("lo", i64), ("hi", i64)]


class Q1Consts(C.Structure):
    _fields_ = [("disc_lo", i64), ("disc_hi", i64), ("qty_lo", i64), ("qty_hi", i64),
                ("price_lo", i64), ("price_hi", i64)]


class ColRef(C.Structure):
    _fields_ = [("comp", vp), ("page_off", vp), ("page_csize", vp), ("page_codec", vp)]


class Dim(C.Structure):
    _fields_ = [("key", vp), ("n", u64), ("preds", vp), ("n_preds", u32), ("group_attr", vp)]


NP_TYPE = {1: np.int8, 2: np.uint8, 3: np.int16, 4: np.uint16, 5: np.int32, 6: np.uint32, 7: np.int64}
TYPE_OF = {np.dtype(v): k for k, v in NP_TYPE.items()}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = C.CDLL(LIB_PATH)
        protos = {
            "oracle_ssb_golden_run": (C.c_int, [u32, C.POINTER(SsbGolden)]),
            "oracle_ssb_sizes": (None, [u32] + [C.POINTER(u64)] * 5),
            "oracle_ssb_generate": (u64, [u32, C.POINTER(SsbCols)]),
            "oracle_splitmix64": (u64, [u64]),
            "oracle_uniform": (u32, [u64, u32, u64, u32, u32]),
            "oracle_gen_q1_fact": (None, [u64, u64, u64, vp, u32, vp, vp, vp, vp]),
            "oracle_gen_date_dim": (None, [u32, vp, vp, vp]),
            "oracle_gen_star_fact": (None, [u64, u64, u64, u64, vp, u32, u32, u32, u32, vp, vp, vp, vp, vp]),
            "oracle_gen_trip": (None, [u64, u64, u64, u64, vp, vp, vp, vp, vp]),
            "oracle_scan_groupby": (i64, [u64, vp, vp, vp, u32, vp, u32, C.c_int32, u32, u64, vp, vp, vp]),
            "oracle_gen_uniform_i32": (None, [u64, u32, u64, C.c_int32, C.c_int32, vp]),
            "oracle_lz4_decode": (i64, [vp, u64, vp, u64]),
            "oracle_lz4i_decode": (i64, [vp, u64, vp, u64]),
            "oracle_dbp_decode": (i64, [vp, u64, vp, u64]),
            "oracle_rle_decode": (i64, [vp, u64, vp, u64]),
            "oracle_page_decode": (i64, [u32, vp, u64, vp, u64]),
            "oracle_decode_column": (C.c_int, [vp, vp, vp, vp, u64, u32, u32, u32, vp]),
            "oracle_q1_plan": (i64, [vp, u64, vp, u32, vp, vp, vp, vp, u64, C.POINTER(Q1Consts), u32,
                                     C.POINTER(u64)]),
            "oracle_q1_compressed": (i64, [vp, u64, vp, u32, vp, u64, u32, u32, u64, u64,
                                           C.POINTER(Q1Consts), u32, C.POINTER(u64)]),
            "oracle_decode_timing": (C.c_int, [vp, vp, u32, u64, u32, u32, u32, C.POINTER(u64),
                                               C.POINTER(C.c_double)]),
            "oracle_star_groupby": (i64, [vp, u32, vp, vp, u64, u32, u64, vp, vp, vp]),
            "oracle_groupby": (i64, [vp, u32, vp, u64, u64, vp, vp, vp]),
            "oracle_minmax_check_i64": (None, [vp, vp, u64, i64, i64, vp]),
            "oracle_minmax_build_i32": (None, [vp, u64, u64, vp, vp]),
            "oracle_eqhist_build": (None, [vp, u64, u64, u64, vp, vp, vp]),
            "oracle_eqhist_check": (None, [vp, vp, vp, u64, u64, i64, i64, vp]),
            "oracle_bloom_params": (None, [C.c_double, u64, C.POINTER(u64), C.POINTER(u64)]),
            "oracle_bloom_build": (None, [vp, u64, u64, u64, u64, vp]),
            "oracle_bloom_check": (None, [vp, u64, u64, u64, i64, vp]),
        }
        for name, (res, args) in protos.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def ssb_golden(sf):
    g = SsbGolden()
    assert load().oracle_ssb_golden_run(sf, C.byref(g)) == 0
    return dict(rows=g.lineorder_rows, q1=list(g.q1), q2_sum=list(g.q2_sum), q2_groups=list(g.q2_groups),
                q34_sum=list(g.q34_sum), q34_groups=list(g.q34_groups))


def ssb_sizes(sf):
    v = [u64() for _ in range(5)]
    load().oracle_ssb_sizes(sf, *[C.byref(x) for x in v])
    return dict(orders=v[0].value, parts=v[1].value, suppliers=v[2].value, customers=v[3].value,
                dates=v[4].value)


def ssb_generate(sf):
    """Columns of the SSB generator (lineorder + the four dimensions): (lo, d, p, s, c) dicts."""
    z = ssb_sizes(sf)
    cap = z["orders"] * 7
    u32c = lambda: np.empty(cap, np.uint32)
    lo = dict(orderdate=np.empty(cap, np.int32), discount=np.empty(cap, np.uint8),
              quantity=np.empty(cap, np.uint8), extendedprice=u32c(), partkey=u32c(), suppkey=u32c(),
              custkey=u32c(), revenue=u32c(), supplycost=u32c())
    d = dict(key=np.empty(2555, np.int32), year=np.empty(2555, np.int16),
             yearmonthnum=np.empty(2555, np.int32), weeknuminyear=np.empty(2555, np.int16))
    p = dict(category=np.empty(z["parts"], np.int16), brand1=np.empty(z["parts"], np.int16),
             mfgr=np.empty(z["parts"], np.uint8))
    sup = dict(region=np.empty(z["suppliers"], np.uint8), nation=np.empty(z["suppliers"], np.uint8),
               city=np.empty(z["suppliers"], np.int16))
    cus = dict(region=np.empty(z["customers"], np.uint8), nation=np.empty(z["customers"], np.uint8),
               city=np.empty(z["customers"], np.int16))
    tabs = dict(lo=lo, d=d, p=p, s=sup, c=cus)
    cols = SsbCols(*[_p(tabs[t][k]) for _, t, k in SSB_COLS])
    n = load().oracle_ssb_generate(sf, C.byref(cols))
    assert n > 0
    lo = {k: v[:n].copy() for k, v in lo.items()}
    return lo, d, p, sup, cus


def uniform(seed, col, row, lo, hi):
    return load().oracle_uniform(seed, col, row, lo, hi)


def gen_date_dim(n_days):
    key = np.empty(n_days, np.int32)
    year = np.empty(n_days, np.int16)
    month = np.empty(n_days, np.uint8)
    load().oracle_gen_date_dim(n_days, _p(key), _p(year), _p(month))
    return key, year, month


def gen_q1_fact(seed, n, date_keys, row0=0):
    date_keys = np.ascontiguousarray(date_keys, dtype=np.int32)
    cols = (np.empty(n, np.int32), np.empty(n, np.uint8), np.empty(n, np.uint8), np.empty(n, np.int32))
    load().oracle_gen_q1_fact(seed, row0, n, _p(date_keys), date_keys.size, *[_p(c) for c in cols])
    return cols


def gen_star_fact(seed, n, n_total, date_keys, n_cust, n_supp, n_part, row0=0):
    date_keys = np.ascontiguousarray(date_keys, dtype=np.int32)
    cols = tuple(np.empty(n, np.int32) for _ in range(5))
    load().oracle_gen_star_fact(seed, row0, n, n_total, _p(date_keys), date_keys.size, n_cust, n_supp, n_part,
                                *[_p(c) for c in cols])
    return cols


def gen_trip(seed, n, n_total=None, row0=0):
    n_total = n if n_total is None else n_total
    cols = (np.empty(n, np.int64), np.empty(n, np.uint8), np.empty(n, np.uint16), np.empty(n, np.int32),
            np.empty(n, np.int32))
    load().oracle_gen_trip(seed, row0, n, n_total, *[_p(c) for c in cols])
    return cols


def gen_uniform_i32(seed, col, n, lo, hi):
    out = np.empty(n, np.int32)
    load().oracle_gen_uniform_i32(seed, col, n, lo, hi, _p(out))
    return out


def lz4_decode(block, n_out):
    src = np.frombuffer(bytes(block), dtype=np.uint8) if len(block) else np.zeros(1, np.uint8)
    dst = np.zeros(max(n_out, 1), np.uint8)
    r = load().oracle_lz4_decode(_p(src), len(block), _p(dst), n_out)
    return None if r < 0 else dst[:r].tobytes()


def page_decode(codec, block, n_out):
    src = np.frombuffer(bytes(block), dtype=np.uint8) if len(block) else np.zeros(1, np.uint8)
    dst = np.zeros(max(n_out, 1), np.uint8)
    r = load().oracle_page_decode(codec, _p(src), len(block), _p(dst), n_out)
    return r, dst


def decode_column(arrs):
    """Decode a whole host column (dict from HostColumn.arrays()) into zero-padded chunk bytes."""
    d = arrs["desc"]
    out = np.zeros(d.n_chunks * d.chunk_rows * d.elem, np.uint8)
    r = load().oracle_decode_column(_p(arrs["comp"]), _p(arrs["page_off"]), _p(arrs["page_csize"]),
                                    _p(arrs["page_codec"]), d.n_rows, d.elem, d.page_rows,
                                    d.chunk_rows, _p(out))
    return None if r != 0 else out


def make_preds(preds):
    """preds: list of (numpy column, lo, hi) -> (ctypes array, keepalive)."""
    arr = (Pred * max(len(preds), 1))()
    keep = []
    for i, (col, lo, hi) in enumerate(preds):
        col = np.ascontiguousarray(col)
        keep.append(col)
        arr[i] = Pred(col.ctypes.data, TYPE_OF[col.dtype], lo, hi)
    return arr, keep


def q1_plan(dim_key, dim_preds, fact_date, disc, qty, price, disc_rng, qty_rng,
            price_rng=(-2**63, 2**63 - 1), threads=1):
    preds, keep = make_preds(dim_preds)
    c = Q1Consts(disc_rng[0], disc_rng[1], qty_rng[0], qty_rng[1], price_rng[0], price_rng[1])
    m = u64()
    dim_key = np.ascontiguousarray(dim_key, np.int32)
    fact_date = np.ascontiguousarray(fact_date, np.int32)
    price = np.ascontiguousarray(price).view(np.int32) if price.dtype == np.uint32 else np.ascontiguousarray(price, np.int32)
    s = load().oracle_q1_plan(_p(dim_key), dim_key.size, C.cast(preds, vp), len(dim_preds), _p(fact_date),
                              _p(np.ascontiguousarray(disc, np.uint8)), _p(np.ascontiguousarray(qty, np.uint8)),
                              _p(price), fact_date.size, C.byref(c), threads, C.byref(m))
    return s, m.value


def q1_compressed(dim_key, dim_preds, col_arrays, n_fact, page_rows, chunk_rows, chunk_begin, chunk_end,
                  disc_rng, qty_rng, price_rng=(-2**63, 2**63 - 1), threads=1):
    """CPU path: per-chunk decode of the four paged columns (date, disc, qty, price) + Q1 plan."""
    preds, keep = make_preds(dim_preds)
    refs = (ColRef * 4)()
    for i, a in enumerate(col_arrays):
        refs[i] = ColRef(a["comp"].ctypes.data, a["page_off"].ctypes.data, a["page_csize"].ctypes.data,
                         a["page_codec"].ctypes.data)
    c = Q1Consts(disc_rng[0], disc_rng[1], qty_rng[0], qty_rng[1], price_rng[0], price_rng[1])
    m = u64()
    dim_key = np.ascontiguousarray(dim_key, np.int32)
    s = load().oracle_q1_compressed(_p(dim_key), dim_key.size, C.cast(preds, vp), len(dim_preds),
                                    C.cast(refs, vp), n_fact, page_rows, chunk_rows, chunk_begin, chunk_end,
                                    C.byref(c), threads, C.byref(m))
    return s, m.value


def decode_timing(col_arrays, elems, n_rows, page_rows, use_liblz4=False, threads=1):
    """Decodes every page of the row-aligned columns (oracle decoders, or liblz4 for LZ4 and
    LZ4I pages); returns (status, checksum, seconds). status -2: liblz4 is not installed.
    col_arrays: HostColumn.arrays() dicts, or (comp, page_off, page_csize, page_codec) addresses."""
    refs = (ColRef * len(col_arrays))()
    for i, a in enumerate(col_arrays):
        if isinstance(a, dict):
            refs[i] = ColRef(a["comp"].ctypes.data, a["page_off"].ctypes.data, a["page_csize"].ctypes.data,
                             a["page_codec"].ctypes.data)
        else:  # (comp, page_off, page_csize, page_codec) addresses
            refs[i] = ColRef(*a)
    el = np.ascontiguousarray(elems, np.uint32)
    cs, sec = u64(), C.c_double()
    r = load().oracle_decode_timing(C.cast(refs, vp), _p(el), len(col_arrays), n_rows, page_rows,
                                    int(bool(use_liblz4)), threads, C.byref(cs), C.byref(sec))
    return r, cs.value, sec.value


def decode_checksum(columns):
    """The checksum decode_timing reports, computed from the raw values: column c's bytes, zero
    padded to 8, as 64-bit words w_g summed with weights 2 * (g + c * 2^40) + 1 (mod 2^64)."""
    total = np.uint64(0)
    with np.errstate(over="ignore"):
        for c, v in enumerate(columns):
            b = np.ascontiguousarray(v).view(np.uint8)
            pad = (-b.size) % 8
            if pad:
                b = np.concatenate([b, np.zeros(pad, np.uint8)])
            w = b.view(np.uint64)
            g = np.arange(w.size, dtype=np.uint64) + np.uint64(c << 40)
            total += np.sum(w * (np.uint64(2) * g + np.uint64(1)), dtype=np.uint64)
    return int(total)


def star_groupby(dims, fks, val, threads=1, max_groups=1 << 22):
    """dims: list of (key int32 array, [(col, lo, hi)...], group_attr int32 array or None)."""
    keep = []
    darr = (Dim * len(dims))()
    ng = 0
    for i, (key, preds, gattr) in enumerate(dims):
        key = np.ascontiguousarray(key, np.int32)
        pa, pk = make_preds(preds)
        keep += [key, pa, pk]
        g = None
        if gattr is not None:
            g = np.ascontiguousarray(gattr, np.int32)
            keep.append(g)
            ng += 1
        darr[i] = Dim(key.ctypes.data, key.size, C.cast(pa, vp), len(preds), g.ctypes.data if g is not None else None)
    fk_arrs = [np.ascontiguousarray(f, np.int32) for f in fks]
    fk_ptrs = (vp * len(fks))(*[f.ctypes.data for f in fk_arrs])
    val = np.ascontiguousarray(val, np.int32)
    keys_out = np.zeros(max_groups * max(ng, 1), np.int32)
    sums = np.zeros(max_groups, np.int64)
    counts = np.zeros(max_groups, np.uint64)
    n = load().oracle_star_groupby(C.cast(darr, vp), len(dims), C.cast(fk_ptrs, vp), _p(val), val.size,
                                   threads, max_groups, _p(keys_out), _p(sums), _p(counts))
    if n < 0:
        return None
    return keys_out[:n * ng].reshape(n, ng), sums[:n], counts[:n]


def groupby(keys, val, max_groups=1 << 22):
    ks = [np.ascontiguousarray(k, np.int32) for k in keys]
    kp = (vp * max(len(ks), 1))(*[k.ctypes.data for k in ks])
    val = np.ascontiguousarray(val, np.int64)
    keys_out = np.zeros(max_groups * max(len(ks), 1), np.int32)
    sums = np.zeros(max_groups, np.int64)
    counts = np.zeros(max_groups, np.uint64)
    n = load().oracle_groupby(C.cast(kp, vp), len(ks), _p(val), val.size, max_groups, _p(keys_out),
                              _p(sums), _p(counts))
    if n < 0:
        return None
    return keys_out[:n * len(ks)].reshape(n, len(ks)), sums[:n], counts[:n]


class ScanPred(C.Structure):
    _fields_ = [("col", u32), ("lo", i64), ("hi", i64)]


class ScanKey(C.Structure):
    _fields_ = [("col", u32), ("base", i64), ("div", u32), ("card", u32), ("out_lo", C.c_int32)]


TYPE_IDS = {np.dtype(np.int8): 1, np.dtype(np.uint8): 2, np.dtype(np.int16): 3, np.dtype(np.uint16): 4,
            np.dtype(np.int32): 5, np.dtype(np.uint32): 6, np.dtype(np.int64): 7}


def scan_groupby(cols, preds, keys, val_col, threads=1, max_groups=1 << 24):
    """cols: list of numpy columns; preds: [(col, lo, hi)]; keys: [(col, base, div, card, out_lo)];
    val_col: index or -1. Returns (keys[n, k], sums, counts) or None if a key is out of range."""
    cs = [np.ascontiguousarray(c) for c in cols]
    cp = (vp * len(cs))(*[c.ctypes.data for c in cs])
    ty = np.array([TYPE_IDS[c.dtype] for c in cs], np.uint32)
    pa = (ScanPred * max(len(preds), 1))(*[ScanPred(*p) for p in preds])
    ka = (ScanKey * max(len(keys), 1))(*[ScanKey(*k) for k in keys])
    nk = max(len(keys), 1)
    dom = 1
    for k in keys:
        dom *= k[3]
    max_groups = min(max_groups, dom)
    keys_out = np.zeros(max_groups * nk, np.int32)
    sums = np.zeros(max_groups, np.int64)
    counts = np.zeros(max_groups, np.uint64)
    n = load().oracle_scan_groupby(cs[0].size, C.cast(cp, vp), _p(ty), C.cast(pa, vp), len(preds), C.cast(ka, vp),
                                   len(keys), val_col, threads, max_groups, _p(keys_out), _p(sums), _p(counts))
    if n < 0:
        return None
    return keys_out[:n * len(keys)].reshape(n, len(keys)), sums[:n], counts[:n]


def minmax_check(mins, maxs, lo, hi):
    mins = np.ascontiguousarray(mins, np.int64)
    maxs = np.ascontiguousarray(maxs, np.int64)
    out = np.zeros(mins.size, np.uint16)
    load().oracle_minmax_check_i64(_p(mins), _p(maxs), mins.size, lo, hi, _p(out))
    return out


def minmax_build(col, chunk_rows):
    col = np.ascontiguousarray(col, np.int32)
    nc = (col.size + chunk_rows - 1) // chunk_rows
    mins = np.zeros(nc, np.int64)
    maxs = np.zeros(nc, np.int64)
    load().oracle_minmax_build_i32(_p(col), col.size, chunk_rows, _p(mins), _p(maxs))
    return mins, maxs


# ---- chunk metadata: EqHistogram / BloomMeta (core/metadata.cuh:225-537)
def eqhist_build(values, chunk_rows, bucket_num):
    """(mins, maxs, hist[chunks, buckets]) of the int column values (EqHistogram + fill_hist)."""
    v = np.ascontiguousarray(values, np.int64)
    nc = (v.size + chunk_rows - 1) // chunk_rows
    mins, maxs = np.zeros(nc, np.int64), np.zeros(nc, np.int64)
    hist = np.zeros(nc * bucket_num, np.uint64)
    load().oracle_eqhist_build(_p(v), v.size, chunk_rows, bucket_num, _p(mins), _p(maxs), _p(hist))
    return mins, maxs, hist.reshape(nc, bucket_num)


def eqhist_check(mins, maxs, hist, lo, hi):
    """check_hist flags of [lo, hi]."""
    hist = np.ascontiguousarray(hist, np.uint64)
    out = np.zeros(mins.size, np.uint16)
    load().oracle_eqhist_check(_p(np.ascontiguousarray(mins, np.int64)), _p(np.ascontiguousarray(maxs, np.int64)),
                               _p(hist), mins.size, hist.shape[1], lo, hi, _p(out))
    return out


def bloom_params(p, m):
    """(n, k) of a (p, m) BloomMeta."""
    n, k = u64(), u64()
    load().oracle_bloom_params(p, m, C.byref(n), C.byref(k))
    return n.value, k.value


def bloom_build(values, chunk_rows, m, k):
    """Per-chunk m-byte filters (as uint32 words) of the int column values (fill_bloom)."""
    v = np.ascontiguousarray(values, np.int64)
    nc = (v.size + chunk_rows - 1) // chunk_rows
    state = np.zeros(nc * m // 4, np.uint32)
    load().oracle_bloom_build(_p(v), v.size, chunk_rows, m, k, _p(state))
    return state


def bloom_check(state, n_chunks, m, k, value):
    """check_bloom flags of one value."""
    out = np.zeros(n_chunks, np.uint16)
    load().oracle_bloom_check(_p(np.ascontiguousarray(state, np.uint32)), n_chunks, m, k, value, _p(out))
    return out
