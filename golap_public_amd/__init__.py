"""ctypes binding of the C ABI in include/golap_hip.h (libgolap_hip.so, built for gfx950).

This module is plumbing for tests and the benchmark driver: every call goes straight to the
native library, which runs the HIP kernels. There is no Python or CPU fallback for any compute
entry point; if the library is missing, load() raises.
"""
import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgolap_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "golap_hip.h")

OK, EINVAL, ENOMEM, EHIP, ECORRUPT, ECAPACITY, EBUDGET, ENODEV = 0, -1, -2, -3, -4, -5, -6, -7
I8, U8, I16, U16, I32, U32, I64 = 1, 2, 3, 4, 5, 6, 7
CODEC_RAW, CODEC_LZ4, CODEC_DBP, CODEC_RLE, CODEC_FOR, CODEC_AUTO = 0, 1, 2, 3, 4, 255
PAGE_LITERAL_ONLY = 0x80

NP_TYPE = {I8: np.int8, U8: np.uint8, I16: np.int16, U16: np.uint16, I32: np.int32,
           U32: np.uint32, I64: np.int64}
TYPE_OF = {np.dtype(v): k for k, v in NP_TYPE.items()}

vp = C.c_void_p
u64 = C.c_uint64
u32 = C.c_uint32
i64 = C.c_int64


class EncodeOpts(C.Structure):
    _fields_ = [("codec", u32), ("page_rows", u32), ("chunk_rows", u32),
                ("lz4_min_match", u32), ("threads", u32), ("lz4_unit", u32)]


class ColumnDesc(C.Structure):
    _fields_ = [("n_rows", u64), ("n_pages", u64), ("n_chunks", u64), ("comp_bytes", u64),
                ("payload_bytes", u64), ("uncomp_bytes", u64), ("type", u32), ("elem", u32),
                ("page_rows", u32), ("chunk_rows", u32), ("comp", vp), ("page_off", vp),
                ("page_csize", vp), ("page_codec", vp), ("chunk_min", vp), ("chunk_max", vp)]


class Pred(C.Structure):
    _fields_ = [("dev_col", vp), ("type", u32), ("lo", i64), ("hi", i64)]


class Q1Plan(C.Structure):
    _fields_ = [("date", vp), ("disc", vp), ("qty", vp), ("price", vp), ("date_map", vp),
                ("disc_lo", i64), ("disc_hi", i64), ("qty_lo", i64), ("qty_hi", i64),
                ("price_lo", i64), ("price_hi", i64), ("chunk_begin", u64), ("chunk_end", u64),
                ("dev_chunk_mask", vp)]


class GroupKey(C.Structure):
    _fields_ = [("dev_col", vp), ("type", u32), ("lo", C.c_int32), ("card", u32)]


class StarDim(C.Structure):
    _fields_ = [("fk", vp), ("map", vp), ("group_lo", C.c_int32), ("group_card", u32)]


class StarPlan(C.Structure):
    _fields_ = [("dims", StarDim * 4), ("n_dims", u32), ("value", vp), ("chunk_begin", u64),
                ("chunk_end", u64), ("dev_chunk_mask", vp), ("agg_mode", u32), ("lds_group_slots", u32),
                ("value_minus", vp)]


class StarResult(C.Structure):
    _fields_ = [("n_groups", u64), ("rows_matched", u64), ("total_sum", i64), ("chunks_run", u64),
                ("chunks_pruned", u64), ("comp_bytes", u64), ("kernel_ms", C.c_double),
                ("driver_dim", u32), ("lds_group_slots", u32)]


AGG_AUTO, AGG_LDS, AGG_GLOBAL = 0, 1, 2


class ScanPred(C.Structure):
    _fields_ = [("col", u32), ("lo", i64), ("hi", i64)]


class ScanKey(C.Structure):
    _fields_ = [("col", u32), ("base", i64), ("div", u32), ("card", u32), ("out_lo", C.c_int32)]


class ScanPlan(C.Structure):
    _fields_ = [("cols", vp * 8), ("n_cols", u32), ("preds", ScanPred * 4), ("n_preds", u32),
                ("keys", ScanKey * 3), ("n_keys", u32), ("value_col", C.c_int32), ("chunk_begin", u64),
                ("chunk_end", u64), ("dev_chunk_mask", vp), ("agg_mode", u32), ("zone_prune", u32),
                ("lds_group_slots", u32)]


class StarQuery(C.Structure):
    _fields_ = [("plan", StarPlan), ("zone_col", vp), ("zone_lo", i64), ("zone_hi", i64)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", u64), ("max_batch", u32), ("slots", u32), ("emit_threads", u32),
                ("batch_window_us", u32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", u64), ("completed", u64), ("refused", u64), ("batches", u64),
                ("peak_admitted_bytes", u64), ("mean_batch", C.c_double), ("device_ms", C.c_double),
                ("comp_bytes", u64), ("shared_comp_bytes", u64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LoadStats(C.Structure):
    _fields_ = [("bytes", u64), ("segments", u64), ("direct_io", u32), ("read_ms", C.c_double),
                ("total_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ChunkMetaDesc(C.Structure):
    _fields_ = [("n_chunks", u64), ("buckets", u32), ("bloom_bytes", u64), ("bloom_k", u64), ("bloom_n", u64),
                ("mins", vp), ("maxs", vp), ("hist", vp), ("bloom_state", vp)]


class BatchConfig(C.Structure):
    _fields_ = [("pages_per_item", u32), ("lds_group_slots", u32), ("lds_cutover", u32), ("filter_lds", u32),
                ("compact_tables", u32), ("max_blocks_per_cu", u32), ("shared_filter", u32),
                ("rows_per_lane", u32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BatchRunInfo(C.Structure):
    _fields_ = [("ms", C.c_double), ("items", u64), ("pages", u64), ("comp_bytes", u64),
                ("shared_comp_bytes", u64)]


class Q1Result(C.Structure):
    _fields_ = [("sum", i64), ("matches", u64), ("chunks_run", u64), ("chunks_pruned", u64),
                ("comp_bytes", u64), ("kernel_ms", C.c_double)]


class HostPred(C.Structure):
    _fields_ = [("col", vp), ("type", u32), ("lo", i64), ("hi", i64)]


class CoexecStats(C.Structure):
    _fields_ = [("rounds", u64), ("splits", u64), ("emitted", u64), ("host_ms", C.c_double), ("device_ms", C.c_double),
                ("total_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Q1FilePlan(C.Structure):
    _fields_ = [("paths", C.c_char_p * 4), ("date_map", vp), ("disc_lo", i64), ("disc_hi", i64), ("qty_lo", i64),
                ("qty_hi", i64), ("price_lo", i64), ("price_hi", i64), ("host_chunk_mask", vp),
                ("host_chunk_mask_len", u64), ("window_chunks", u32), ("zone_prune", u32)]


class StarFilePlan(C.Structure):
    _fields_ = [("fk_paths", C.c_char_p * 4), ("maps", vp * 4), ("group_lo", C.c_int32 * 4), ("group_card", u32 * 4),
                ("n_dims", u32), ("value_path", C.c_char_p), ("host_chunk_mask", vp), ("host_chunk_mask_len", u64),
                ("window_chunks", u32), ("agg_mode", u32)]


class StreamStats(C.Structure):
    _fields_ = [("windows", u64), ("windows_skipped", u64), ("bytes_read", u64), ("bytes_copied", u64),
                ("direct_io", u32), ("window_chunks", u32), ("read_ms", C.c_double), ("total_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_PROTOS = {
    "golap_api_version": (C.c_int, []),
    "golap_status_string": (C.c_char_p, [C.c_int]),
    "golap_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "golap_ctx_destroy": (C.c_int, [vp]),
    "golap_last_error": (C.c_char_p, [vp]),
    "golap_sync": (C.c_int, [vp]),
    "golap_device_info": (C.c_int, [vp, C.POINTER(C.c_int), C.c_char_p, C.c_int]),
    "golap_dev_alloc": (C.c_int, [vp, u64, C.POINTER(vp)]),
    "golap_dev_free": (C.c_int, [vp, vp]),
    "golap_h2d": (C.c_int, [vp, vp, vp, u64]),
    "golap_d2h": (C.c_int, [vp, vp, vp, u64]),
    "golap_memset": (C.c_int, [vp, vp, C.c_int, u64]),
    "golap_d2d": (C.c_int, [vp, vp, vp, u64]),
    "golap_read_probe": (C.c_int, [vp, vp, u64, u32, C.POINTER(C.c_double)]),
    "golap_hashmap_keys": (C.c_int, [vp, vp, vp, u64, C.POINTER(u64)]),
    "golap_host_column_save": (C.c_int, [vp, C.c_char_p]),
    "golap_host_column_load": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
    "golap_encode_device": (C.c_int, [vp, vp, u64, u32, C.c_void_p, C.POINTER(vp), C.POINTER(C.c_double)]),
    "golap_column_load_file": (C.c_int, [vp, C.c_char_p, u64, u32, C.POINTER(vp), C.c_void_p]),
    "golap_eqhist_build": (C.c_int, [vp, vp, u32, u64, u64, u32, C.POINTER(vp)]),
    "golap_eqhist_check": (C.c_int, [vp, vp, i64, i64, vp]),
    "golap_bloommeta_build": (C.c_int, [vp, vp, u32, u64, u64, C.c_double, u64, C.POINTER(vp)]),
    "golap_bloommeta_check": (C.c_int, [vp, vp, i64, vp]),
    "golap_chunkmeta_info": (C.c_int, [vp, C.c_void_p]),
    "golap_chunkmeta_free": (C.c_int, [vp, vp]),
    "golap_minmax_check_points": (C.c_int, [vp, vp, vp, u64, vp]),
    "golap_device_allocated": (u64, [vp]),
    "golap_encode_opts_default": (None, [C.POINTER(EncodeOpts)]),
    "golap_encode": (C.c_int, [vp, u64, u32, C.POINTER(EncodeOpts), C.POINTER(vp)]),
    "golap_host_column_assemble": (C.c_int, [u32, u64, u32, u32, C.POINTER(vp),
                                             C.POINTER(u32), C.POINTER(C.c_uint8), C.POINTER(vp)]),
    "golap_host_column_free": (C.c_int, [vp]),
    "golap_host_column_info": (C.c_int, [vp, C.POINTER(ColumnDesc)]),
    "golap_column_upload": (C.c_int, [vp, vp, C.POINTER(vp)]),
    "golap_column_free": (C.c_int, [vp, vp]),
    "golap_column_info": (C.c_int, [vp, C.POINTER(ColumnDesc)]),
    "golap_prepare_compressed_device": (C.c_int, [vp, vp, u64, u32, C.POINTER(EncodeOpts),
                                                  C.POINTER(vp), C.POINTER(C.c_double)]),
    "golap_decompress_chunks": (C.c_int, [vp, vp, u64, u64, vp, C.POINTER(C.c_double)]),
    "golap_hashmap_create": (C.c_int, [vp, u64, vp, u32, C.POINTER(vp)]),
    "golap_hashmap_build": (C.c_int, [vp, vp, u64, C.POINTER(Pred), u32, vp]),
    "golap_hashmap_probe": (C.c_int, [vp, vp, vp, u32, u64, vp]),
    "golap_hash_join_count": (C.c_int, [vp, vp, vp, u32, u64, C.POINTER(u64)]),
    "golap_hashmap_filled": (C.c_int, [vp, vp, C.POINTER(u64)]),
    "golap_hashmap_free": (C.c_int, [vp, vp]),
    "golap_minmax_check": (C.c_int, [vp, vp, i64, i64, vp]),
    "golap_combine_and": (C.c_int, [vp, vp, vp, u64]),
    "golap_q1_run": (C.c_int, [vp, C.POINTER(Q1Plan), C.POINTER(Q1Result)]),
    "golap_star_run_files": (C.c_int, [vp, C.POINTER(StarFilePlan), C.POINTER(StarResult), C.POINTER(StreamStats),
                                       vp, vp, vp, u64]),
    "golap_q1_run_files": (C.c_int, [vp, C.POINTER(Q1FilePlan), C.POINTER(Q1Result), C.POINTER(StreamStats)]),
    "golap_group_by": (C.c_int, [vp, C.POINTER(GroupKey), u32, vp, u32, u64, u32, C.POINTER(u64), vp, vp, vp,
                                 u64]),
    "golap_star_run": (C.c_int, [vp, C.POINTER(StarPlan), C.POINTER(StarResult), vp, vp, vp, u64]),
    "golap_host_hashmap_create": (C.c_int, [u64, vp, C.POINTER(vp)]),
    "golap_host_hashmap_build": (C.c_int, [vp, u64, C.c_void_p, u32, vp, u32]),
    "golap_host_hashmap_probe": (C.c_int, [vp, vp, u64, vp, C.POINTER(u64)]),
    "golap_host_hashmap_filled": (C.c_int, [vp, C.POINTER(u64)]),
    "golap_host_hashmap_free": (C.c_int, [vp]),
    "golap_batch_results": (C.c_int, [vp, vp, vp, vp, vp, vp, u64]),
    "golap_star_run_hostdim": (C.c_int, [vp, C.POINTER(StarPlan), u32, vp, u32, C.POINTER(StarResult),
                                         C.POINTER(CoexecStats), vp, vp, vp, u64]),
    "golap_scan_run": (C.c_int, [vp, C.POINTER(ScanPlan), C.POINTER(StarResult), vp, vp, vp, u64]),
    "golap_scan_run_noalloc": (C.c_int, [vp, C.POINTER(ScanPlan), C.POINTER(StarResult), vp, vp, vp, u64]),
    "golap_batch_create_star": (C.c_int, [vp, C.POINTER(StarQuery), u32, C.POINTER(vp)]),
    "golap_batch_create_scan": (C.c_int, [vp, C.POINTER(ScanPlan), u32, C.POINTER(vp)]),
    "golap_batch_config_get": (C.c_int, [vp, C.POINTER(BatchConfig)]),
    "golap_batch_config_set": (C.c_int, [vp, C.POINTER(BatchConfig)]),
    "golap_batch_run": (C.c_int, [vp, u32, C.POINTER(BatchRunInfo)]),
    "golap_batch_tune": (C.c_int, [vp, u32, C.POINTER(BatchConfig), C.POINTER(C.c_double)]),
    "golap_batch_result": (C.c_int, [vp, u32, C.POINTER(StarResult), vp, vp, vp, u64]),
    "golap_batch_free": (C.c_int, [vp]),
    "golap_server_create": (C.c_int, [vp, C.POINTER(ServerOpts), C.POINTER(vp)]),
    "golap_server_submit_star": (C.c_int, [vp, C.POINTER(StarQuery), C.POINTER(u64)]),
    "golap_server_submit_scan": (C.c_int, [vp, C.POINTER(ScanPlan), C.POINTER(u64)]),
    "golap_server_wait": (C.c_int, [vp, u64, C.POINTER(StarResult), vp, vp, vp, u64, C.POINTER(C.c_double)]),
    "golap_server_get_stats": (C.c_int, [vp, C.POINTER(ServerStats)]),
    "golap_server_destroy": (C.c_int, [vp]),
    "golap_gen_trip": (C.c_int, [u64, u64, u64, u64, vp, vp, vp, vp, vp, u32]),
    "golap_gen_uniform_i32": (C.c_int, [u64, u32, u64, C.c_int32, C.c_int32, vp, u32]),
    "golap_gen_date_dim": (C.c_int, [u32, vp, vp, vp]),
    "golap_gen_q1_fact": (C.c_int, [u64, u64, u64, vp, u32, vp, vp, vp, vp, u32]),
    "golap_stream_create": (C.c_int, [vp, C.POINTER(vp)]),
    "golap_stream_destroy": (C.c_int, [vp]),
    "golap_stream_sync": (C.c_int, [vp]),
    "golap_event_create": (C.c_int, [vp, C.POINTER(vp)]),
    "golap_event_destroy": (C.c_int, [vp]),
    "golap_event_record": (C.c_int, [vp, vp]),
    "golap_stream_wait_event": (C.c_int, [vp, vp, vp]),
    "golap_event_sync": (C.c_int, [vp]),
    "golap_stream_copy": (C.c_int, [vp, vp, vp, vp, u64]),
    "golap_decompress_chunks_async": (C.c_int, [vp, vp, u64, u64, vp, vp]),
    "golap_hashagg_create": (C.c_int, [vp, C.POINTER(GroupKey), u32, u32, C.POINTER(vp)]),
    "golap_hashagg_add": (C.c_int, [vp, vp, vp, u32, u64, vp]),
    "golap_hashagg_results": (C.c_int, [vp, C.POINTER(u64), vp, vp, vp, u64]),
    "golap_hashagg_clear": (C.c_int, [vp]),
    "golap_hashagg_destroy": (C.c_int, [vp]),
    "golap_prejoin_minmax": (C.c_int, [vp, vp, vp, C.POINTER(vp)]),
    "golap_chunkmeta_minmax_check": (C.c_int, [vp, vp, i64, i64, vp]),
}

_lib = None


def load():
    """Load libgolap_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with python -m golap_public_amd.build")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class GolapError(RuntimeError):
    def __init__(self, status, msg=""):
        super().__init__(f"status {status} ({load().golap_status_string(status).decode()}) {msg}")
        self.status = status


def _domain(dims):
    n = 1
    for d in dims:
        if d[3]:
            n *= int(d[3])
    return n


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def encode_opts(codec=CODEC_LZ4, page_rows=4096, chunk_rows=262144, lz4_min_match=4, threads=0, lz4_unit=1):
    o = EncodeOpts()
    load().golap_encode_opts_default(C.byref(o))
    o.codec, o.page_rows, o.chunk_rows, o.lz4_min_match, o.threads, o.lz4_unit = (
        codec, page_rows, chunk_rows, lz4_min_match, threads, lz4_unit)
    return o


class HostHashMap:
    """HostHashMap (core/host_structs.hpp:9-78) over a host int32 key column."""

    def __init__(self, keys, ht_size=None):
        self.keys = np.ascontiguousarray(keys, dtype=np.int32)
        h = vp()
        rc = load().golap_host_hashmap_create(ht_size or max(1, self.keys.size), _ptr(self.keys), C.byref(h))
        if rc:
            raise GolapError(rc, "host_hashmap_create")
        self.h = h.value

    def build(self, preds=(), payload=None, threads=0, n_rows=None):
        """preds: [(host numpy column, lo, hi)]; payload: int32 per row or None (store rids)."""
        n = self.keys.size if n_rows is None else n_rows
        cols = [np.ascontiguousarray(c) for c, _, _ in preds]
        if n > self.keys.size or any(c.size < n for c in cols) or (payload is not None and len(payload) < n):
            raise ValueError("HostHashMap.build: keys, predicate columns and payload need n_rows entries")
        arr = (HostPred * max(1, len(preds)))()
        for i, (c, (_, lo, hi)) in enumerate(zip(cols, preds)):
            arr[i] = HostPred(c.ctypes.data, TYPE_OF[c.dtype], lo, hi)
        pay = None if payload is None else np.ascontiguousarray(payload, dtype=np.int32)
        rc = load().golap_host_hashmap_build(self.h, n, C.cast(arr, C.c_void_p), len(preds),
                                             None if pay is None else _ptr(pay), threads)
        if rc:
            raise GolapError(rc, "host_hashmap_build")

    def probe(self, keys):
        keys = np.ascontiguousarray(keys, dtype=np.int32)
        out = np.zeros(keys.size, np.uint64)
        hits = u64()
        rc = load().golap_host_hashmap_probe(self.h, _ptr(keys), keys.size, _ptr(out), C.byref(hits))
        if rc:
            raise GolapError(rc, "host_hashmap_probe")
        return out, hits.value

    @property
    def filled(self):
        f = u64()
        load().golap_host_hashmap_filled(self.h, C.byref(f))
        return f.value

    def free(self):
        if self.h:
            load().golap_host_hashmap_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HostColumn:
    """An encoded column in host memory (CompInfo + BlockInfo + MinMaxMeta)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle)

    @classmethod
    def encode(cls, values, **kw):
        values = np.ascontiguousarray(values)
        t = TYPE_OF[values.dtype]
        out = vp()
        st = load().golap_encode(_ptr(values), values.size, t, C.byref(encode_opts(**kw)), C.byref(out))
        if st != OK:
            raise GolapError(st, "encode")
        return cls(out.value)

    @classmethod
    def assemble(cls, type_, n_rows, page_rows, chunk_rows, pages, codecs):
        n = len(pages)
        bufs = [np.frombuffer(p, dtype=np.uint8) if len(p) else np.zeros(1, np.uint8) for p in pages]
        ptrs = (vp * max(n, 1))(*[b.ctypes.data for b in bufs])
        sizes = (u32 * max(n, 1))(*[len(p) for p in pages])
        cods = (C.c_uint8 * max(n, 1))(*codecs)
        out = vp()
        st = load().golap_host_column_assemble(type_, n_rows, page_rows, chunk_rows, ptrs, sizes, cods,
                                               C.byref(out))
        if st != OK:
            raise GolapError(st, "assemble")
        return cls(out.value)

    def save(self, path):
        """Writes the column file (header, catalog, 4 KiB-aligned page data)."""
        st = load().golap_host_column_save(self.h, str(path).encode())
        if st != OK:
            raise GolapError(st, "host_column_save")

    @classmethod
    def load_file(cls, path):
        out = vp()
        st = load().golap_host_column_load(str(path).encode(), C.byref(out))
        if st != OK:
            raise GolapError(st, "host_column_load")
        return cls(out.value)

    def desc(self):
        d = ColumnDesc()
        st = load().golap_host_column_info(self.h, C.byref(d))
        if st != OK:
            raise GolapError(st)
        return d

    def arrays(self):
        """Copies of the catalog and the column buffer as numpy arrays."""
        d = self.desc()

        def arr(ptr, n, dt):
            if n == 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dt))),
                                         shape=(n,)).copy()
        return dict(comp=arr(d.comp, d.comp_bytes, np.uint8),
                    page_off=arr(d.page_off, d.n_pages, np.uint64),
                    page_csize=arr(d.page_csize, d.n_pages, np.uint32),
                    page_codec=arr(d.page_codec, d.n_pages, np.uint8),
                    chunk_min=arr(d.chunk_min, d.n_chunks, np.int64),
                    chunk_max=arr(d.chunk_max, d.n_chunks, np.int64),
                    desc=d)

    def free(self):
        if self.h:
            load().golap_host_column_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One device; owns every device object created through it."""

    def __init__(self, device=0):
        self.lib = load()
        h = vp()
        st = self.lib.golap_ctx_create(device, C.byref(h))
        if st != OK:
            raise GolapError(st, "ctx_create")
        self.h = h

    def check(self, st, what=""):
        if st != OK:
            raise GolapError(st, f"{what}: {self.lib.golap_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            self.lib.golap_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self):
        cus = C.c_int()
        name = C.create_string_buffer(128)
        self.check(self.lib.golap_device_info(self.h, C.byref(cus), name, 128))
        return cus.value, name.value.decode()

    # memory
    def alloc(self, nbytes):
        p = vp()
        self.check(self.lib.golap_dev_alloc(self.h, nbytes, C.byref(p)), "alloc")
        return p.value

    def free(self, p):
        self.check(self.lib.golap_dev_free(self.h, p), "free")

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        self.check(self.lib.golap_h2d(self.h, p, _ptr(arr), arr.nbytes), "h2d")
        return p

    def h2d(self, p, arr):
        arr = np.ascontiguousarray(arr)
        self.check(self.lib.golap_h2d(self.h, p, _ptr(arr), arr.nbytes), "h2d")

    def to_host(self, p, n, dtype):
        out = np.empty(n, dtype=dtype)
        self.check(self.lib.golap_d2h(self.h, _ptr(out), p, out.nbytes), "d2h")
        return out

    def memset(self, p, value, nbytes):
        self.check(self.lib.golap_memset(self.h, p, value, nbytes), "memset")

    def copy(self, dst, src, nbytes):
        self.check(self.lib.golap_d2d(self.h, dst, src, nbytes), "d2d")

    def read_probe(self, p, nbytes, blocks_per_cu=2):
        """Device ms of one streaming read of nbytes at device pointer p."""
        ms = C.c_double()
        self.check(self.lib.golap_read_probe(self.h, p, nbytes, blocks_per_cu, C.byref(ms)), "read_probe")
        return ms.value

    def sync(self):
        self.check(self.lib.golap_sync(self.h), "sync")

    # columns
    def upload(self, hcol):
        out = vp()
        self.check(self.lib.golap_column_upload(self.h, hcol.h, C.byref(out)), "upload")
        return out.value

    def prepare_compressed_device(self, values, **kw):
        values = np.ascontiguousarray(values)
        out = vp()
        ms = C.c_double()
        self.check(self.lib.golap_prepare_compressed_device(
            self.h, _ptr(values), values.size, TYPE_OF[values.dtype], C.byref(encode_opts(**kw)),
            C.byref(out), C.byref(ms)), "prepare_compressed_device")
        return out.value, ms.value

    def encode_device(self, dev_values, n_rows, type_, **kw):
        """prepare_compressed on the device: values already in HBM coded into a new device
        column, byte-identical to HostColumn.encode + upload. Returns (column, kernel_ms)."""
        out, ms = vp(), C.c_double()
        self.check(self.lib.golap_encode_device(self.h, dev_values, n_rows, type_, C.byref(encode_opts(**kw)),
                                                C.byref(out), C.byref(ms)), "encode_device")
        return out.value, ms.value

    def load_file(self, path, segment_bytes=0, n_buffers=0):
        """SSD -> HBM: the column file's pages through pinned buffers into a new device column.
        Returns (column handle, LoadStats)."""
        out, st = vp(), LoadStats()
        self.check(self.lib.golap_column_load_file(self.h, str(path).encode(), segment_bytes, n_buffers, C.byref(out),
                                                   C.byref(st)), "column_load_file")
        return out.value, st

    def column_desc(self, col):
        d = ColumnDesc()
        self.check(self.lib.golap_column_info(col, C.byref(d)))
        return d

    def free_column(self, col):
        self.check(self.lib.golap_column_free(self.h, col), "column_free")

    def decompress(self, col, chunk_begin=0, chunk_end=None):
        """Decode chunks into a fresh device buffer and return it as a host numpy array."""
        d = self.column_desc(col)
        if chunk_end is None:
            chunk_end = d.n_chunks
        nbytes = (chunk_end - chunk_begin) * d.chunk_rows * d.elem
        buf = self.alloc(max(nbytes, 16))
        try:
            ms = C.c_double()
            self.check(self.lib.golap_decompress_chunks(self.h, col, chunk_begin, chunk_end, buf,
                                                        C.byref(ms)), "decompress")
            out = self.to_host(buf, nbytes, np.uint8)
        finally:
            self.free(buf)
        return out, ms.value

    # streams, events, asynchronous decode, device-resident group-by
    def stream(self):
        """A stream of its own (golap_stream_create); release with free_stream."""
        h = vp()
        self.check(self.lib.golap_stream_create(self.h, C.byref(h)), "stream_create")
        return h.value

    def free_stream(self, s):
        self.lib.golap_stream_destroy(s)

    def stream_sync(self, s):
        self.check(self.lib.golap_stream_sync(s), "stream_sync")

    def event(self):
        h = vp()
        self.check(self.lib.golap_event_create(self.h, C.byref(h)), "event_create")
        return h.value

    def free_event(self, e):
        self.lib.golap_event_destroy(e)

    def event_record(self, e, stream=None):
        self.check(self.lib.golap_event_record(e, stream), "event_record")

    def stream_wait_event(self, stream, e):
        self.check(self.lib.golap_stream_wait_event(self.h, stream, e), "stream_wait_event")

    def decompress_async(self, col, dev_out, chunk_begin=0, chunk_end=None, stream=None):
        if chunk_end is None:
            chunk_end = self.column_desc(col).n_chunks
        self.check(self.lib.golap_decompress_chunks_async(self.h, col, chunk_begin, chunk_end, dev_out, stream),
                   "decompress_async")

    def hashagg(self, keys, agg_mode=AGG_AUTO):
        """keys: list of (type, lo, card). Returns a golap_hashagg handle."""
        arr = (GroupKey * len(keys))()
        for i, (t, lo, card) in enumerate(keys):
            arr[i] = GroupKey(None, t, lo, card)
        h = vp()
        self.check(self.lib.golap_hashagg_create(self.h, arr, len(keys), agg_mode, C.byref(h)), "hashagg_create")
        return h.value

    def hashagg_add(self, h, dev_keys, dev_val, val_type, n, stream=None):
        ks = (vp * len(dev_keys))(*dev_keys)
        self.check(self.lib.golap_hashagg_add(h, ks, dev_val, val_type, n, stream), "hashagg_add")

    def hashagg_results(self, h, n_keys, max_groups):
        ng = u64()
        k = np.zeros(max_groups * n_keys, np.int32)
        sums = np.zeros(max_groups, np.int64)
        cnts = np.zeros(max_groups, np.uint64)
        self.check(self.lib.golap_hashagg_results(h, C.byref(ng), _ptr(k), _ptr(sums), _ptr(cnts), max_groups),
                   "hashagg_results")
        n = ng.value
        return k[:n * n_keys].reshape(n, n_keys), sums[:n], cnts[:n]

    def free_hashagg(self, h):
        self.lib.golap_hashagg_destroy(h)

    # hash join
    def hashmap(self, ht_size, dev_keys, key_type):
        out = vp()
        self.check(self.lib.golap_hashmap_create(self.h, ht_size, dev_keys, key_type, C.byref(out)),
                   "hashmap_create")
        return out.value

    def hashmap_build(self, m, n_rows, preds=(), dev_payload=None):
        arr = (Pred * max(len(preds), 1))(*[Pred(*p) for p in preds])
        st = self.lib.golap_hashmap_build(self.h, m, n_rows, arr, len(preds), dev_payload)
        self.check(st, "hashmap_build")

    def hashmap_filled(self, m):
        v = u64()
        self.check(self.lib.golap_hashmap_filled(self.h, m, C.byref(v)))
        return v.value

    def hashmap_probe(self, m, dev_keys, key_type, n, dev_out):
        self.check(self.lib.golap_hashmap_probe(self.h, m, dev_keys, key_type, n, dev_out), "probe")

    def hash_join_count(self, m, dev_keys, key_type, n):
        v = u64()
        self.check(self.lib.golap_hash_join_count(self.h, m, dev_keys, key_type, n, C.byref(v)), "count")
        return v.value

    def free_hashmap(self, m):
        self.check(self.lib.golap_hashmap_free(self.h, m), "hashmap_free")

    # zone maps
    def minmax_check(self, col, lo, hi, dev_flags):
        self.check(self.lib.golap_minmax_check(self.h, col, lo, hi, dev_flags), "minmax_check")

    def combine_and(self, dev_a, dev_b, n):
        self.check(self.lib.golap_combine_and(self.h, dev_a, dev_b, n), "combine_and")

    def eqhist_build(self, dev_values, type_, n_rows, chunk_rows, buckets):
        """EqHistogram per chunk of a device column; returns a handle."""
        h = vp()
        self.check(self.lib.golap_eqhist_build(self.h, dev_values, type_, n_rows, chunk_rows, buckets, C.byref(h)),
                   "eqhist_build")
        return h.value

    def eqhist_check(self, h, lo, hi, dev_flags):
        self.check(self.lib.golap_eqhist_check(self.h, h, lo, hi, dev_flags), "eqhist_check")

    def bloommeta_build(self, dev_values, type_, n_rows, chunk_rows, p, m_bytes):
        """BloomMeta per chunk of a device column; returns a handle."""
        h = vp()
        self.check(self.lib.golap_bloommeta_build(self.h, dev_values, type_, n_rows, chunk_rows, p, m_bytes,
                                                  C.byref(h)), "bloommeta_build")
        return h.value

    def bloommeta_check(self, h, value, dev_flags):
        self.check(self.lib.golap_bloommeta_check(self.h, h, value, dev_flags), "bloommeta_check")

    def chunkmeta_desc(self, h):
        d = ChunkMetaDesc()
        self.check(self.lib.golap_chunkmeta_info(h, C.byref(d)), "chunkmeta_info")
        return d

    def prejoin_minmax(self, fk_col, dim_map):
        """Per-chunk min/max of the dimension attribute (the map's payload) fk_col's rows join to."""
        h = vp()
        self.check(self.lib.golap_prejoin_minmax(self.h, fk_col, dim_map, C.byref(h)), "prejoin_minmax")
        return h.value

    def chunkmeta_minmax_check(self, h, lo, hi, dev_flags):
        self.check(self.lib.golap_chunkmeta_minmax_check(self.h, h, lo, hi, dev_flags), "chunkmeta_minmax_check")

    def chunkmeta_free(self, h):
        self.check(self.lib.golap_chunkmeta_free(self.h, h), "chunkmeta_free")

    def hashmap_keys(self, m, dev_keys, cap):
        """collect_hashmap_keys: the map's keys, ascending, into dev_keys; returns their count."""
        n = u64()
        self.check(self.lib.golap_hashmap_keys(self.h, m, dev_keys, cap, C.byref(n)), "hashmap_keys")
        return n.value

    def minmax_check_points(self, col, dev_keys, n, dev_flags):
        """check_mmmeta_multiple_pts: chunk flags = some key in the chunk's [min, max]."""
        self.check(self.lib.golap_minmax_check_points(self.h, col, dev_keys, n, dev_flags), "minmax_check_points")

    # group-by
    def group_by(self, keys, dev_val, val_type, n_rows, agg_mode=AGG_AUTO, max_groups=1 << 20):
        """keys: list of (dev_col, type, lo, card). Returns (keys[n, k], sums, counts)."""
        arr = (GroupKey * len(keys))(*[GroupKey(*k) for k in keys])
        ng = u64()
        kout = np.zeros(max_groups * len(keys), np.int32)
        sums = np.zeros(max_groups, np.int64)
        cnts = np.zeros(max_groups, np.uint64)
        self.check(self.lib.golap_group_by(self.h, arr, len(keys), dev_val, val_type, n_rows, agg_mode,
                                           C.byref(ng), _ptr(kout), _ptr(sums), _ptr(cnts), max_groups), "group_by")
        n = ng.value
        return kout[:n * len(keys)].reshape(n, len(keys)), sums[:n], cnts[:n]

    def star_plan(self, dims, value, chunk_begin=0, chunk_end=None, dev_chunk_mask=None, agg_mode=AGG_AUTO,
                  lds_group_slots=0, value_minus=None):
        if chunk_end is None:
            chunk_end = self.column_desc(value).n_chunks
        p = StarPlan()
        for i, d in enumerate(dims):
            p.dims[i] = StarDim(*d)
        p.n_dims = len(dims)
        p.value, p.chunk_begin, p.chunk_end = value, chunk_begin, chunk_end
        p.dev_chunk_mask, p.agg_mode, p.lds_group_slots = dev_chunk_mask, agg_mode, lds_group_slots
        p.value_minus = value_minus
        return p

    def star_run(self, dims, value, chunk_begin=0, chunk_end=None, dev_chunk_mask=None, agg_mode=AGG_AUTO,
                 lds_group_slots=0, max_groups=None, value_minus=None):
        """dims: list of (fk column, map, group_lo, group_card). Returns (result, keys, sums, counts).
        max_groups defaults to the group domain (the product of the group cards). value_minus: sum
        value - value_minus."""
        p = self.star_plan(dims, value, chunk_begin, chunk_end, dev_chunk_mask, agg_mode, lds_group_slots,
                           value_minus)
        na = sum(1 for d in dims if d[3])
        if max_groups is None:
            max_groups = _domain(dims)
        kout = np.zeros(max_groups * max(na, 1), np.int32)
        sums = np.zeros(max_groups, np.int64)
        cnts = np.zeros(max_groups, np.uint64)
        r = StarResult()
        self.check(self.lib.golap_star_run(self.h, C.byref(p), C.byref(r), _ptr(kout), _ptr(sums), _ptr(cnts),
                                           max_groups), "star_run")
        n = r.n_groups
        return r, kout[:n * na].reshape(n, na), sums[:n], cnts[:n]

    def star_run_files(self, dims, value_path, host_chunk_mask=None, window_chunks=0, agg_mode=AGG_AUTO,
                       max_groups=None):
        """A star plan streamed from column files: dims = [(fk file, map, group_lo, group_card)].
        Returns (result, keys, sums, counts, StreamStats)."""
        p = StarFilePlan()
        paths = [str(d[0]).encode() for d in dims]
        for i, (fk, m, glo, gcard) in enumerate(dims):
            p.fk_paths[i], p.maps[i], p.group_lo[i], p.group_card[i] = paths[i], m, glo, gcard
        p.n_dims = len(dims)
        vpath = str(value_path).encode()
        p.value_path = vpath
        mask = None
        if host_chunk_mask is not None:
            mask = np.ascontiguousarray(host_chunk_mask, dtype=np.uint16)
            p.host_chunk_mask, p.host_chunk_mask_len = mask.ctypes.data, mask.size
        p.window_chunks, p.agg_mode = window_chunks, agg_mode
        na = sum(1 for d in dims if d[3])
        if max_groups is None:
            max_groups = _domain(dims)
        kout = np.zeros(max_groups * max(na, 1), np.int32)
        sums = np.zeros(max_groups, np.int64)
        cnts = np.zeros(max_groups, np.uint64)
        r, st = StarResult(), StreamStats()
        self.check(self.lib.golap_star_run_files(self.h, C.byref(p), C.byref(r), C.byref(st), _ptr(kout), _ptr(sums),
                                                 _ptr(cnts), max_groups), "star_run_files")
        n = r.n_groups
        return r, kout[:n * na].reshape(n, na), sums[:n], cnts[:n], st

    def star_run_hostdim(self, dims, value, host_dim, hmap, threads=0, chunk_begin=0, chunk_end=None,
                         dev_chunk_mask=None, max_groups=None):
        """A star plan with dims[host_dim] joined on the host through hmap (a HostHashMap; that
        dim's map entry is ignored). Returns (result, keys, sums, counts, CoexecStats)."""
        dims = [(d[0], d[1] if i != host_dim else None, d[2], d[3]) for i, d in enumerate(dims)]
        p = self.star_plan(dims, value, chunk_begin, chunk_end, dev_chunk_mask, AGG_GLOBAL, 0)
        na = sum(1 for d in dims if d[3])
        if max_groups is None:
            max_groups = _domain(dims)
        kout = np.zeros(max_groups * max(na, 1), np.int32)
        sums = np.zeros(max_groups, np.int64)
        cnts = np.zeros(max_groups, np.uint64)
        r, st = StarResult(), CoexecStats()
        self.check(self.lib.golap_star_run_hostdim(self.h, C.byref(p), host_dim, hmap.h, threads, C.byref(r),
                                                   C.byref(st), _ptr(kout), _ptr(sums), _ptr(cnts), max_groups),
                   "star_run_hostdim")
        n = r.n_groups
        return r, kout[:n * na].reshape(n, na), sums[:n], cnts[:n], st

    def scan_plan(self, cols, preds=(), keys=(), value_col=-1, chunk_begin=0, chunk_end=None, dev_chunk_mask=None,
                  agg_mode=AGG_AUTO, zone_prune=1, lds_group_slots=0):
        """cols: device columns; preds: [(col, lo, hi)]; keys: [(col, base, div, card, out_lo)]."""
        if chunk_end is None:
            chunk_end = self.column_desc(cols[0]).n_chunks
        p = ScanPlan()
        for i, c in enumerate(cols):
            p.cols[i] = c
        p.n_cols = len(cols)
        for i, pr in enumerate(preds):
            p.preds[i] = ScanPred(*pr)
        p.n_preds = len(preds)
        for i, k in enumerate(keys):
            p.keys[i] = ScanKey(*k)
        p.n_keys = len(keys)
        p.value_col, p.chunk_begin, p.chunk_end = value_col, chunk_begin, chunk_end
        p.dev_chunk_mask, p.agg_mode, p.zone_prune, p.lds_group_slots = dev_chunk_mask, agg_mode, zone_prune, \
            lds_group_slots
        return p

    def scan_run(self, cols, preds=(), keys=(), value_col=-1, max_groups=None, noalloc=False, **kw):
        """Returns (result, keys[n, k], sums, counts); max_groups defaults to the group domain.
        noalloc: through golap_scan_run_noalloc (no device allocation)."""
        p = self.scan_plan(cols, preds, keys, value_col, **kw)
        if max_groups is None:
            max_groups = 1
            for k in keys:
                max_groups *= int(k[3])
        nk = len(keys)
        kout = np.zeros(max_groups * max(nk, 1), np.int32)
        sums = np.zeros(max_groups, np.int64)
        cnts = np.zeros(max_groups, np.uint64)
        r = StarResult()
        fn = self.lib.golap_scan_run_noalloc if noalloc else self.lib.golap_scan_run
        self.check(fn(self.h, C.byref(p), C.byref(r), _ptr(kout), _ptr(sums), _ptr(cnts), max_groups), "scan_run")
        n = r.n_groups
        return r, kout[:n * nk].reshape(n, nk), sums[:n], cnts[:n]

    def star_batch(self, queries):
        """queries: list of dicts with star_plan() arguments plus optional zone=(column, lo, hi)."""
        return Batch(self, "star", queries)

    def server(self, **opts):
        """Serving loop over this context (golap_server_*); opts: ServerOpts fields."""
        return Server(self, **opts)

    def scan_batch(self, plans):
        """plans: list of dicts with scan_plan() arguments."""
        return Batch(self, "scan", plans)

    # plans
    def q1_run(self, date, disc, qty, price, date_map, disc_rng, qty_rng, price_rng=(-2**31, 2**31 - 1),
               chunk_begin=0, chunk_end=None, dev_chunk_mask=None):
        if chunk_end is None:
            chunk_end = self.column_desc(date).n_chunks
        p = Q1Plan(date, disc, qty, price, date_map, disc_rng[0], disc_rng[1], qty_rng[0],
                   qty_rng[1], price_rng[0], price_rng[1], chunk_begin, chunk_end, dev_chunk_mask)
        r = Q1Result()
        self.check(self.lib.golap_q1_run(self.h, C.byref(p), C.byref(r)), "q1_run")
        return r

    def q1_run_files(self, paths, date_map, disc_rng, qty_rng, price_rng=(-2**31, 2**31 - 1),
                     host_chunk_mask=None, window_chunks=0, zone_prune=True):
        """Q1 streamed from four column files (date, disc, qty, price) in windows of chunks;
        host_chunk_mask: optional uint16 per chunk (0 = pruned, never read).
        Returns (Q1Result, StreamStats)."""
        if len(paths) != 4:
            raise ValueError("q1_run_files: four column files (date, disc, qty, price)")
        mask = None
        if host_chunk_mask is not None:
            mask = np.ascontiguousarray(host_chunk_mask, dtype=np.uint16)
        p = Q1FilePlan((C.c_char_p * 4)(*[str(x).encode() for x in paths]), date_map, disc_rng[0], disc_rng[1],
                       qty_rng[0], qty_rng[1], price_rng[0], price_rng[1],
                       None if mask is None else mask.ctypes.data, 0 if mask is None else mask.size,
                       window_chunks, 1 if zone_prune else 0)
        r, st = Q1Result(), StreamStats()
        self.check(self.lib.golap_q1_run_files(self.h, C.byref(p), C.byref(r), C.byref(st)), "q1_run_files")
        return r, st


class Batch:
    """Many star or scan plans over one fact table in one persistent launch (golap_batch_*)."""

    def __init__(self, ctx, kind, queries):
        self.ctx, self.h, self.kind = ctx, None, kind
        self.n_attrs, self.domains = [], []
        h = vp()
        if kind == "star":
            arr = (StarQuery * len(queries))()
            for i, q in enumerate(queries):
                q = dict(q)
                zone = q.pop("zone", None)
                arr[i].plan = ctx.star_plan(**q)
                if zone is not None:
                    arr[i].zone_col, arr[i].zone_lo, arr[i].zone_hi = zone
                self.n_attrs.append(sum(1 for d in q["dims"] if d[3]))
                self.domains.append(_domain(q["dims"]))
            st = ctx.lib.golap_batch_create_star(ctx.h, arr, len(queries), C.byref(h))
        else:
            arr = (ScanPlan * len(queries))()
            for i, q in enumerate(queries):
                arr[i] = ctx.scan_plan(**q)
                keys = q.get("keys", ())
                self.n_attrs.append(len(keys))
                d = 1
                for k in keys:
                    d *= int(k[3])
                self.domains.append(d)
            st = ctx.lib.golap_batch_create_scan(ctx.h, arr, len(queries), C.byref(h))
        ctx.check(st, "batch_create")
        self.h = h.value
        self.n = len(queries)

    def config(self):
        c = BatchConfig()
        self.ctx.check(self.ctx.lib.golap_batch_config_get(self.h, C.byref(c)), "batch_config_get")
        return c

    def set_config(self, **kw):
        c = self.config()
        for k, v in kw.items():
            setattr(c, k, v)
        self.ctx.check(self.ctx.lib.golap_batch_config_set(self.h, C.byref(c)), "batch_config_set")

    def run(self, use_graph=True):
        info = BatchRunInfo()
        self.ctx.check(self.ctx.lib.golap_batch_run(self.h, int(use_graph), C.byref(info)), "batch_run")
        return info

    def tune(self, reps=3):
        c, ms = BatchConfig(), C.c_double()
        self.ctx.check(self.ctx.lib.golap_batch_tune(self.h, reps, C.byref(c), C.byref(ms)), "batch_tune")
        return c, ms.value

    def result(self, q, max_groups=None):
        na = self.n_attrs[q]
        if max_groups is None:
            max_groups = self.domains[q]
        kout = np.zeros(max_groups * max(na, 1), np.int32)
        sums = np.zeros(max_groups, np.int64)
        cnts = np.zeros(max_groups, np.uint64)
        r = StarResult()
        self.ctx.check(self.ctx.lib.golap_batch_result(self.h, q, C.byref(r), _ptr(kout), _ptr(sums),
                                                       _ptr(cnts), max_groups), "batch_result")
        n = r.n_groups
        return r, kout[:n * na].reshape(n, na), sums[:n], cnts[:n]

    def results(self, max_total=None):
        """Every query's (result, keys, sums, counts) of the last run, from one call."""
        nq = len(self.n_attrs)
        if max_total is None:
            max_total = max(1, min(sum(self.domains), 1 << 22))
        na_max = max([1] + list(self.n_attrs))
        while True:
            res = (StarResult * nq)()
            off = np.zeros(nq + 1, np.uint64)
            keys = np.empty(max_total * na_max, np.int32)
            sums = np.empty(max_total, np.int64)
            cnts = np.empty(max_total, np.uint64)
            rc = self.ctx.lib.golap_batch_results(self.h, C.cast(res, C.c_void_p), _ptr(off), _ptr(keys), _ptr(sums),
                                                  _ptr(cnts), max_total)
            if rc == ECAPACITY and int(off[-1]) > max_total:
                max_total = int(off[-1])
                continue
            self.ctx.check(rc, "batch_results")
            break
        out, ko = [], 0
        for q in range(nq):
            g0, g1, na = int(off[q]), int(off[q + 1]), self.n_attrs[q]
            n = g1 - g0
            out.append((res[q], keys[ko:ko + n * na].reshape(n, na), sums[g0:g1], cnts[g0:g1]))
            ko += n * na
        return out

    def free(self):
        if self.h:
            self.ctx.check(self.ctx.lib.golap_batch_free(self.h), "batch_free")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()


class Server:
    """Request loop with HBM-budget admission and continuous batching (golap_server_*)."""

    def __init__(self, ctx, **opts):
        self.ctx, self.h = ctx, None
        o = ServerOpts()
        for k, v in opts.items():
            setattr(o, k, v)
        h = vp()
        ctx.check(ctx.lib.golap_server_create(ctx.h, C.byref(o), C.byref(h)), "server_create")
        self.h = h.value
        self.shape = {}
        self._lock = threading.Lock()

    def submit_scan(self, **plan):
        p = self.ctx.scan_plan(**plan)
        t = u64()
        self.ctx.check(self.ctx.lib.golap_server_submit_scan(self.h, C.byref(p), C.byref(t)), "server_submit")
        keys = plan.get("keys", ())
        d = 1
        for k in keys:
            d *= int(k[3])
        with self._lock:
            self.shape[t.value] = (len(keys), d)
        return t.value

    def submit_star(self, zone=None, **query):
        q = StarQuery()
        q.plan = self.ctx.star_plan(**query)
        if zone is not None:
            q.zone_col, q.zone_lo, q.zone_hi = zone
        t = u64()
        self.ctx.check(self.ctx.lib.golap_server_submit_star(self.h, C.byref(q), C.byref(t)), "server_submit")
        with self._lock:
            self.shape[t.value] = (sum(1 for d in query["dims"] if d[3]), _domain(query["dims"]))
        return t.value

    def wait(self, ticket):
        """(result, keys, sums, counts, latency_ms)."""
        with self._lock:
            if ticket not in self.shape:
                raise GolapError(EINVAL, "server_wait: unknown ticket or one already waited on")
            na, dom = self.shape.pop(ticket)
        kout = np.zeros(dom * max(na, 1), np.int32)
        sums = np.zeros(dom, np.int64)
        cnts = np.zeros(dom, np.uint64)
        r, lat = StarResult(), C.c_double()
        self.ctx.check(self.ctx.lib.golap_server_wait(self.h, ticket, C.byref(r), _ptr(kout), _ptr(sums), _ptr(cnts),
                                                      dom, C.byref(lat)), "server_wait")
        n = r.n_groups
        return r, kout[:n * na].reshape(n, na), sums[:n], cnts[:n], lat.value

    def stats(self):
        st = ServerStats()
        self.ctx.check(self.ctx.lib.golap_server_get_stats(self.h, C.byref(st)), "server_stats")
        return st.as_dict()

    def close(self):
        if self.h:
            self.ctx.check(self.ctx.lib.golap_server_destroy(self.h), "server_destroy")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def gen_date_dim(n_days):
    key = np.empty(n_days, np.int32)
    year = np.empty(n_days, np.int16)
    month = np.empty(n_days, np.uint8)
    load().golap_gen_date_dim(n_days, _ptr(key), _ptr(year), _ptr(month))
    return key, year, month


def gen_q1_fact(seed, n, date_keys, row0=0, threads=0):
    """Seeded Q1 fact columns (date_key, discount, quantity, price) for rows [row0, row0+n)."""
    date_keys = np.ascontiguousarray(date_keys, dtype=np.int32)
    cols = (np.empty(n, np.int32), np.empty(n, np.uint8), np.empty(n, np.uint8), np.empty(n, np.int32))
    st = load().golap_gen_q1_fact(seed, row0, n, _ptr(date_keys), date_keys.size,
                                  *[_ptr(c) for c in cols], threads)
    if st != OK:
        raise GolapError(st, "gen_q1_fact")
    return cols


def gen_star_fact(seed, n, n_total, date_keys, n_cust, n_supp, n_part, row0=0, threads=0):
    """Star fact columns (date_key sorted, cust, supp, part keys, revenue) for rows [row0, row0+n)."""
    date_keys = np.ascontiguousarray(date_keys, dtype=np.int32)
    cols = tuple(np.empty(n, np.int32) for _ in range(5))
    st = load().golap_gen_star_fact(seed, row0, n, n_total, _ptr(date_keys), date_keys.size, n_cust, n_supp,
                                    n_part, *[_ptr(c) for c in cols], threads)
    if st != OK:
        raise GolapError(st, "gen_star_fact")
    return cols


TRIP_T0, TRIP_SPAN = 1704067200, 31622400


def gen_trip(seed, n, n_total=None, row0=0, threads=0):
    """Trip-table columns (pickup_ts int64 sorted over 2024, passenger uint8 Zipf(1.2) 1..6, zone
    uint16 1..265, fare_cents int32 lognormal(7.2, 0.6) in [250, 50000], distance_m int32
    lognormal(8, 0.8)) for rows [row0, row0+n) of n_total."""
    n_total = n if n_total is None else n_total
    cols = (np.empty(n, np.int64), np.empty(n, np.uint8), np.empty(n, np.uint16), np.empty(n, np.int32),
            np.empty(n, np.int32))
    st = load().golap_gen_trip(seed, row0, n, n_total, *[_ptr(c) for c in cols], threads)
    if st != OK:
        raise GolapError(st, "gen_trip")
    return cols


def gen_uniform_i32(seed, col, n, lo, hi, threads=0):
    out = np.empty(n, np.int32)
    st = load().golap_gen_uniform_i32(seed, col, n, lo, hi, _ptr(out), threads)
    if st != OK:
        raise GolapError(st, "gen_uniform_i32")
    return out


def header_symbols():
    """Function names declared in include/golap_hip.h."""
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(golap_[a-z_0-9]+)\s*\(",
                                 text, re.M)))
