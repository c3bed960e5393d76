"""ctypes binding of libmirage_hip (include/mirage_hip.h).

Device memory for inputs and outputs is held in torch tensors (plumbing only);
every computation runs in the HIP library.  Loading fails loudly when the
library has not been built: there is no fallback path.
"""
import ctypes
import os

import numpy as np

from . import _build

MH_OK = 0
MH_ERR_ARG = -1
MH_ERR_HIP = -2
MH_ERR_NOMEM = -3
MH_ERR_RANGE = -4
MH_ERR_OVERFLOW = -5
MH_ERR_BADBASE = -6
MH_ERR_STATE = -7
MH_INDEX_LAST_BUCKET_EMPTY = 1
MH_CLASSIFY_ULTRAFAST = 1
MH_CLASSIFY_NO_DIRECT = 2
MH_CLASSIFY_BATCHED = 4
MH_CLASSIFY_GRAPH = 16

READ_RESULT_DTYPE = np.dtype([("highest", "<f4"), ("strand", "<u4"), ("n_top", "<u4"),
                              ("top_off", "<u4"), ("n_cand", "<u4"), ("status", "<u4")])
CAND_DTYPE = np.dtype([("read", "<u4"), ("strand", "<u4"), ("ref", "<u4"), ("hits", "<u4"),
                       ("sum_r", "<u8"), ("sum_c", "<u8"), ("score", "<f4"), ("rel", "<f4"),
                       ("n_seeds", "<u4"), ("pad", "<u4")])
COUNTER_FIELDS = ("queries", "probes", "hits", "entries", "seeds", "cands", "direct", "overflow",
                  "ext_compares")
MH_CLASSIFY_TIMING = 8
MH_CLASSIFY_PREPROBE = 32
MH_READ_BADBASE = 1
MH_READ_TOO_LONG = 2
MH_READ_CAPACITY = 4
MH_NOALLOC_SCRATCH_BYTES = 256 << 20
MH_NOALLOC_MAX_READS = 1 << 20


class IndexInfo(ctypes.Structure):
    _fields_ = [("n_entries", ctypes.c_uint64), ("n_buckets", ctypes.c_uint64),
                ("device_bytes", ctypes.c_uint64), ("n_refs", ctypes.c_uint32),
                ("threshold", ctypes.c_int32), ("empty_hash", ctypes.c_uint32),
                ("pad", ctypes.c_uint32)]


class MhmInfo(ctypes.Structure):
    _fields_ = [("n_buckets", ctypes.c_uint64), ("file_bytes", ctypes.c_uint64),
                ("record_bytes", ctypes.c_uint64), ("ref_bytes", ctypes.c_uint64),
                ("header_bytes", ctypes.c_uint64), ("n_refs", ctypes.c_uint32),
                ("threshold", ctypes.c_int32), ("hash_bytes", ctypes.c_uint32),
                ("streampos_bytes", ctypes.c_uint32)]


class LogThresholds(ctypes.Structure):
    _fields_ = [("high_sensitivity", ctypes.c_float), ("high_precision", ctypes.c_float),
                ("balanced", ctypes.c_float)]


MH_RESULTS_APPEND = 1
MH_RESULTS_ALL_GROUPS = 2


class Counters(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint64) for f in COUNTER_FIELDS]


class Timing(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_float), ("bucket_ms", ctypes.c_float * 8),
                ("big_ms", ctypes.c_float), ("merge_ms", ctypes.c_float),
                ("n_buckets", ctypes.c_uint32), ("probe_ms", ctypes.c_float),
                ("tier_ms", ctypes.c_float * 4)]


class Tile(ctypes.Structure):
    _fields_ = [("max_len", ctypes.c_uint32), ("block", ctypes.c_uint32),
                ("entries_cap", ctypes.c_uint32), ("grid", ctypes.c_uint32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("max_inflight", ctypes.c_uint32),
                ("max_batch_reads", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("pad", ctypes.c_uint32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("request_id", ctypes.c_uint64), ("n_reads", ctypes.c_uint32),
                ("status", ctypes.c_int32), ("submit_s", ctypes.c_double),
                ("admit_s", ctypes.c_double), ("done_s", ctypes.c_double)]


# Every symbol declared in include/mirage_hip.h.
EXPORTS = ("mh_open", "mh_close", "mh_last_error", "mh_stream", "mh_max_read_len",
           "mh_max_batch_reads", "mh_index_build", "mh_index_free", "mh_index_get_info",
           "mh_index_bucket", "mh_classify", "mh_classify_host", "mh_classify_noalloc",
           "mh_get_counters", "mh_get_kernel_counters", "mh_get_timing", "mh_score_seeds",
           "mh_set_tiles", "mh_get_tiles", "mh_set_overflow_tiles", "mh_autotune",
           "mh_server_start", "mh_server_submit", "mh_server_wait", "mh_server_stop",
           "mh_mhm_open", "mh_mhm_close", "mh_mhm_get_info", "mh_mhm_references", "mh_mhm_bucket",
           "mh_index_from_mhm", "mh_index_save_mhm", "mh_write_results", "mh_fasta_parse")

_lib = None


class MirageError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status})")
        self.status = status


def load():
    """Loads libmirage_hip.so; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    path = _build.hip_lib_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() first")
    L = ctypes.CDLL(path)
    P, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32
    sig = {
        "mh_open": (i32, [i32, u32, ctypes.POINTER(P)]),
        "mh_close": (i32, [P]),
        "mh_last_error": (ctypes.c_char_p, [P]),
        "mh_stream": (P, [P]),
        "mh_max_read_len": (u32, [P]),
        "mh_max_batch_reads": (u64, [P]),
        "mh_index_build": (i32, [P, P, P, u32, i32, u32, ctypes.POINTER(P)]),
        "mh_index_free": (i32, [P]),
        "mh_index_get_info": (i32, [P, ctypes.POINTER(IndexInfo)]),
        "mh_index_bucket": (i32, [P, P, u32, P, P, u64, ctypes.POINTER(ctypes.c_int64)]),
        "mh_classify": (i32, [P, P, P, P, u32, u32, P, P, u64, P, u64, P, P]),
        "mh_classify_host": (i32, [P, P, P, P, u32, u32, P, P, u64, P, u64, P, P]),
        "mh_classify_noalloc": (i32, [P, P, P, P, u32, u32, P, P, u64, P, u64, P, P]),
        "mh_get_counters": (i32, [P, ctypes.POINTER(Counters)]),
        "mh_get_timing": (i32, [P, ctypes.POINTER(Timing)]),
        "mh_get_kernel_counters": (i32, [P, P, u32, ctypes.POINTER(u32)]),
        "mh_score_seeds": (i32, [P, P, P, P, P, P, P, P, u32, u32, P, P, P, P, P]),
        "mh_set_tiles": (i32, [P, P, u32]),
        "mh_get_tiles": (i32, [P, P, P, u32, ctypes.POINTER(u32), P]),
        "mh_set_overflow_tiles": (i32, [P, P]),
        "mh_autotune": (i32, [P, P, P, P, u32, u32]),
        "mh_server_start": (i32, [P, P, ctypes.POINTER(ServerConfig), ctypes.POINTER(P)]),
        "mh_server_submit": (i32, [P, P, P, u32, ctypes.POINTER(u64)]),
        "mh_server_wait": (i32, [P, u64, ctypes.c_double, P, P, u64, ctypes.POINTER(u64),
                                 ctypes.POINTER(RequestInfo)]),
        "mh_server_stop": (i32, [P]),
        "mh_mhm_open": (i32, [ctypes.c_char_p, ctypes.POINTER(P)]),
        "mh_mhm_close": (i32, [P]),
        "mh_mhm_get_info": (i32, [P, ctypes.POINTER(MhmInfo)]),
        "mh_mhm_references": (i32, [P, P, P, P, P, P]),
        "mh_mhm_bucket": (i32, [P, u32, u32, P, P, u64, ctypes.POINTER(ctypes.c_int64)]),
        "mh_index_from_mhm": (i32, [P, P, i32, u32, ctypes.POINTER(P)]),
        "mh_index_save_mhm": (i32, [P, P, ctypes.c_char_p, P, P, P, u32, u32]),
        "mh_fasta_parse": (i32, [P, P, u64, P, u64, P, P, u64, P, u64, ctypes.POINTER(u64),
                                 ctypes.POINTER(u64), ctypes.POINTER(u64)]),
        "mh_write_results": (i32, [ctypes.c_char_p, u64, u32, P, P, P, P, P, u64, u32, P, P, P,
                                   ctypes.POINTER(LogThresholds), u32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _torch_sync(t):
    """Waits for torch's current stream on t's device, so that work torch
    queued on t (a copy from the host, say) is complete before the library's
    own stream reads it."""
    import torch
    torch.cuda.current_stream(t.device).synchronize()


def _np_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _dev_ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Handle:
    """One device, one HIP stream (mh_open)."""

    def __init__(self, device=0, max_read_len=0):
        self.lib = load()
        h = ctypes.c_void_p()
        st = self.lib.mh_open(device, max_read_len, ctypes.byref(h))
        if st != MH_OK:
            raise MirageError(st, "mh_open failed")
        self.h = h
        self.device = device

    def check(self, st, what):
        if st != MH_OK:
            raise MirageError(st, f"{what}: {self.lib.mh_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            self.lib.mh_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def max_read_len(self):
        return self.lib.mh_max_read_len(self.h)

    def stream_ptr(self):
        return self.lib.mh_stream(self.h)

    def build_index(self, ref_bytes, ref_off, threshold, last_bucket_empty=True):
        return DeviceIndex(self, ref_bytes, ref_off, threshold, last_bucket_empty)

    def load_mhm(self, mhm, threshold=0, last_bucket_empty=True):
        """Device index from a .mhm file (a path or an open MhmFile),
        decoded in HBM (mh_index_from_mhm)."""
        f = mhm if isinstance(mhm, MhmFile) else MhmFile(mhm)
        try:
            ix = ctypes.c_void_p()
            flags = MH_INDEX_LAST_BUCKET_EMPTY if last_bucket_empty else 0
            self.check(self.lib.mh_index_from_mhm(self.h, f.m, threshold, flags, ctypes.byref(ix)),
                       "mh_index_from_mhm")
            info = f.info()
        finally:
            if f is not mhm:
                f.close()
        return DeviceIndex.wrap(self, ix, threshold or info["threshold"], info["n_refs"])

    def parse_fasta(self, d_text, exact=False):
        """FASTA text in a device tensor (uint8) to device tensors (sequence
        bytes, n+1 offsets, header bytes, n+1 offsets) (mh_fasta_parse).
        Torch's current stream is synchronised first (the library's stream
        does not wait on it).  One call parses into buffers sized from the
        text (sequence bytes never exceed it; headers and records are
        guessed, and a second, exactly sized call follows when the guess is
        short); exact=True asks for the sizes first instead."""
        import torch
        _torch_sync(d_text)
        n = int(d_text.numel())
        dev = d_text.device
        nr, ns, nh = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()

        def call(seq, so, hdr, ho, caps):
            return self.lib.mh_fasta_parse(
                self.h, _dev_ptr(d_text), n, _dev_ptr(seq) if seq is not None else None, caps[0],
                _dev_ptr(so) if so is not None else None,
                _dev_ptr(hdr) if hdr is not None else None, caps[1],
                _dev_ptr(ho) if ho is not None else None, caps[2],
                ctypes.byref(nr), ctypes.byref(ns), ctypes.byref(nh))

        if exact:
            self.check(call(None, None, None, None, (0, 0, 0)), "mh_fasta_parse")
            caps = (ns.value, nh.value, nr.value)
        else:
            caps = (n, n // 2 + 64, n // 32 + 16)
        for attempt in range(2):
            seq = torch.empty(max(caps[0], 1), dtype=torch.uint8, device=dev)
            hdr = torch.empty(max(caps[1], 1), dtype=torch.uint8, device=dev)
            so = torch.empty(caps[2] + 1, dtype=torch.int64, device=dev)
            ho = torch.empty(caps[2] + 1, dtype=torch.int64, device=dev)
            st = call(seq, so, hdr, ho, caps)
            if st == MH_ERR_OVERFLOW and attempt == 0:
                caps = (ns.value, nh.value, nr.value)
                continue
            self.check(st, "mh_fasta_parse")
            break
        r = nr.value
        return seq[:ns.value], so[:r + 1], hdr[:nh.value], ho[:r + 1]

    def set_tiles(self, tiles):
        """Installs a length-tile table (mh_set_tiles); None restores the
        index-sized defaults."""
        if tiles is None:
            self.check(self.lib.mh_set_tiles(self.h, None, 0), "mh_set_tiles")
            return
        arr = (Tile * len(tiles))(*[Tile(*t) for t in tiles])
        self.check(self.lib.mh_set_tiles(self.h, ctypes.cast(arr, ctypes.c_void_p), len(tiles)),
                   "mh_set_tiles")

    def get_tiles(self, index):
        """(length tiles, overflow tiers) a batched call on `index` uses."""
        arr = (Tile * 8)()
        tiers = (Tile * 2)()
        n = ctypes.c_uint32()
        self.check(self.lib.mh_get_tiles(self.h, index.ix, ctypes.cast(arr, ctypes.c_void_p), 8,
                                         ctypes.byref(n), ctypes.cast(tiers, ctypes.c_void_p)),
                   "mh_get_tiles")
        as_tuple = lambda t: (t.max_len, t.block, t.entries_cap, t.grid)  # noqa: E731
        return [as_tuple(arr[i]) for i in range(n.value)], [as_tuple(t) for t in tiers]

    def set_overflow_tiles(self, tiers):
        """The two LDS overflow tiers (mh_set_overflow_tiles); None restores
        the defaults."""
        if tiers is None:
            self.check(self.lib.mh_set_overflow_tiles(self.h, None), "mh_set_overflow_tiles")
            return
        arr = (Tile * 2)(*[Tile(*t) for t in tiers])
        self.check(self.lib.mh_set_overflow_tiles(self.h, ctypes.cast(arr, ctypes.c_void_p)),
                   "mh_set_overflow_tiles")

    def timing(self):
        t = Timing()
        self.check(self.lib.mh_get_timing(self.h, ctypes.byref(t)), "mh_get_timing")
        return {"total_ms": t.total_ms, "bucket_ms": list(t.bucket_ms), "big_ms": t.big_ms,
                "merge_ms": t.merge_ms, "n_buckets": t.n_buckets, "probe_ms": t.probe_ms,
                "tier_ms": list(t.tier_ms)}

    def autotune(self, index, d_reads, d_read_off, n_reads, iters=3):
        """Times tile candidates per length bucket on device-resident reads and
        installs the fastest (mh_autotune).  Returns the tile table."""
        self.check(self.lib.mh_autotune(self.h, index.ix, _dev_ptr(d_reads), _dev_ptr(d_read_off),
                                        n_reads, iters), "mh_autotune")
        return self.get_tiles(index)[0]

    def kernel_counters(self):
        """Per-kernel counters of the last call: rows 0-7 length tiles,
        8-11 the overflow tiers (second chance, middle, HBM, last resort)."""
        arr = (Counters * 12)()
        n = ctypes.c_uint32()
        self.check(self.lib.mh_get_kernel_counters(self.h, ctypes.cast(arr, ctypes.c_void_p), 12,
                                                   ctypes.byref(n)), "mh_get_kernel_counters")
        return [{f: getattr(arr[i], f) for f in COUNTER_FIELDS} for i in range(n.value)]

    def counters(self):
        c = Counters()
        self.check(self.lib.mh_get_counters(self.h, ctypes.byref(c)), "mh_get_counters")
        return {f: getattr(c, f) for f in COUNTER_FIELDS}


class DeviceIndex:
    """Bucketed k-mer index resident in HBM (mh_index_build)."""

    def __init__(self, handle, ref_bytes, ref_off, threshold, last_bucket_empty=True):
        self.handle = handle
        rb = np.ascontiguousarray(ref_bytes, dtype=np.uint8)
        ro = np.ascontiguousarray(ref_off, dtype=np.uint64)
        ix = ctypes.c_void_p()
        flags = MH_INDEX_LAST_BUCKET_EMPTY if last_bucket_empty else 0
        handle.check(handle.lib.mh_index_build(handle.h, _np_ptr(rb), _np_ptr(ro), len(ro) - 1,
                                               threshold, flags, ctypes.byref(ix)),
                     "mh_index_build")
        self.ix = ix
        self.threshold = threshold
        self.n_refs = len(ro) - 1

    @classmethod
    def wrap(cls, handle, ix, threshold, n_refs):
        self = cls.__new__(cls)
        self.handle, self.ix, self.threshold, self.n_refs = handle, ix, threshold, n_refs
        return self

    def save_mhm(self, path, headers=None, taxids=None, hash_bytes=4, streampos_bytes=4):
        """Writes the index as a .mhm v4 file (mh_index_save_mhm).  headers:
        per reference the bytes MirAge stores ('>' name NUL), or None."""
        hb = ho = tx = None
        if headers is not None:
            hb = np.frombuffer(b"".join(headers) or b"\0", np.uint8)
            ho = np.zeros(len(headers) + 1, np.uint64)
            ho[1:] = np.cumsum([len(x) for x in headers])
        if taxids is not None:
            tx = np.ascontiguousarray(taxids, dtype=np.int32)
        self.handle.check(self.handle.lib.mh_index_save_mhm(
            self.handle.h, self.ix, os.fsencode(path), _np_ptr(hb), _np_ptr(ho), _np_ptr(tx),
            hash_bytes, streampos_bytes), "mh_index_save_mhm")

    def free(self):
        if self.ix:
            self.handle.lib.mh_index_free(self.ix)
            self.ix = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def info(self):
        i = IndexInfo()
        self.handle.check(self.handle.lib.mh_index_get_info(self.ix, ctypes.byref(i)),
                          "mh_index_get_info")
        return {f: getattr(i, f) for f, _ in IndexInfo._fields_ if f != "pad"}

    def bucket(self, h):
        n = ctypes.c_int64()
        lib = self.handle.lib
        self.handle.check(lib.mh_index_bucket(self.handle.h, self.ix, h, None, None, 0,
                                              ctypes.byref(n)), "mh_index_bucket")
        if n.value < 0:
            return None
        refs = np.zeros(max(n.value, 1), np.uint32)
        pos = np.zeros(max(n.value, 1), np.uint32)
        self.handle.check(lib.mh_index_bucket(self.handle.h, self.ix, h, _np_ptr(refs),
                                              _np_ptr(pos), n.value, ctypes.byref(n)),
                          "mh_index_bucket")
        return list(zip(refs[:n.value].tolist(), pos[:n.value].tolist()))

    def classify_host(self, reads, read_off, flags=0, top_cap=None, want_cands=False,
                      cand_cap=None):
        """Full path from host arrays (mh_classify_host).  Returns
        (results, top_pool, cands)."""
        lib, h = self.handle.lib, self.handle.h
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        read_off = np.ascontiguousarray(read_off, dtype=np.uint64)
        n = len(read_off) - 1
        top_cap = top_cap or max(64, 16 * n)
        cand_cap = cand_cap or max(64, 16 * n)
        for _ in range(3):
            out = np.zeros(n, READ_RESULT_DTYPE)
            top = np.zeros(top_cap, np.uint32)
            cands = np.zeros(cand_cap, CAND_DTYPE) if want_cands else None
            tn, cn = ctypes.c_uint64(), ctypes.c_uint64()
            st = lib.mh_classify_host(h, self.ix, _np_ptr(reads), _np_ptr(read_off), n, flags,
                                      _np_ptr(out), _np_ptr(top), top_cap, _np_ptr(cands),
                                      cand_cap if want_cands else 0, ctypes.byref(tn),
                                      ctypes.byref(cn))
            if st == MH_ERR_OVERFLOW:
                top_cap = max(top_cap, int(tn.value) + 16)
                cand_cap = max(cand_cap, int(cn.value) + 16)
                continue
            self.handle.check(st, "mh_classify_host")
            if want_cands:
                cands = cands[:cn.value]
            return out, top[:tn.value], cands
        raise MirageError(MH_ERR_OVERFLOW, "output pools kept overflowing")

    def score_seeds(self, queries, query_off, pair_query, pair_ref, seed_off, seeds, flags=0):
        """Scoring only (mh_score_seeds) from host arrays: pair p scores query
        pair_query[p] against reference pair_ref[p] with seeds
        seeds[seed_off[p]:seed_off[p+1]] as (start.ref, start.query, last.ref,
        last.query) rows.  Returns (score, rel, sum_r, sum_c, hits) arrays."""
        import torch
        q = np.ascontiguousarray(queries, dtype=np.uint8)
        qo = np.ascontiguousarray(query_off, dtype=np.uint64)
        pq = np.ascontiguousarray(pair_query, dtype=np.uint32)
        pr = np.ascontiguousarray(pair_ref, dtype=np.uint32)
        so = np.ascontiguousarray(seed_off, dtype=np.uint64)
        sd = np.ascontiguousarray(seeds, dtype=np.uint32).reshape(-1)
        n = len(pq)
        dev = lambda a: torch.from_numpy(a.view(np.uint8).copy() if a.size else  # noqa: E731
                                         np.zeros(8, np.uint8)).cuda()
        d_q, d_qo, d_pq, d_pr, d_so = dev(q), dev(qo), dev(pq), dev(pr), dev(so)
        d_sd = dev(sd) if sd.size else None
        outs = [torch.zeros(max(n, 1) * w, dtype=torch.uint8, device="cuda") for w in (4, 4, 8, 8, 4)]
        _torch_sync(d_q)
        st = self.handle.lib.mh_score_seeds(
            self.handle.h, self.ix, _dev_ptr(d_q), _dev_ptr(d_qo), _dev_ptr(d_pq), _dev_ptr(d_pr),
            _dev_ptr(d_so), _dev_ptr(d_sd), n, flags, *[_dev_ptr(o) for o in outs])
        self.handle.check(st, "mh_score_seeds")
        torch.cuda.synchronize()
        types = (np.float32, np.float32, np.uint64, np.uint64, np.uint32)
        return tuple(o.cpu().numpy().view(t)[:n] for o, t in zip(outs, types))

    def classify_noalloc(self, d_reads, d_read_off, n_reads, d_out, d_top, top_cap, flags=0,
                         d_cands=None, cand_cap=0, wait=True):
        """mh_classify_noalloc: the device path with the library's fixed
        scratch array, no allocation.  Same arguments as classify_device."""
        _torch_sync(d_reads)
        lib, h = self.handle.lib, self.handle.h
        tn, cn = ctypes.c_uint64(), ctypes.c_uint64()
        st = lib.mh_classify_noalloc(h, self.ix, _dev_ptr(d_reads), _dev_ptr(d_read_off), n_reads,
                                     flags, _dev_ptr(d_out), _dev_ptr(d_top), top_cap,
                                     _dev_ptr(d_cands), cand_cap,
                                     ctypes.byref(tn) if wait else None,
                                     ctypes.byref(cn) if wait else None)
        self.handle.check(st, "mh_classify_noalloc")
        return (tn.value, cn.value) if wait else None

    def classify_device(self, d_reads, d_read_off, n_reads, d_out, d_top, top_cap, flags=0,
                        d_cands=None, cand_cap=0, wait=True, sync_inputs=True):
        """Full path on device-resident torch tensors (mh_classify).  The
        library's stream does not wait on torch's: with sync_inputs the
        current torch stream is synchronised first, so tensors just written
        (a copy from the host, say) are complete."""
        if sync_inputs:
            _torch_sync(d_reads)
        lib, h = self.handle.lib, self.handle.h
        tn, cn = ctypes.c_uint64(), ctypes.c_uint64()
        st = lib.mh_classify(h, self.ix, _dev_ptr(d_reads), _dev_ptr(d_read_off), n_reads, flags,
                             _dev_ptr(d_out), _dev_ptr(d_top), top_cap, _dev_ptr(d_cands),
                             cand_cap, ctypes.byref(tn) if wait else None,
                             ctypes.byref(cn) if wait else None)
        self.handle.check(st, "mh_classify")
        return (tn.value, cn.value) if wait else None


def _headers(hs):
    """Concatenated header bytes and n+1 offsets from a list of bytes."""
    b = np.frombuffer(b"".join(hs) or b"\0", np.uint8)
    off = np.zeros(len(hs) + 1, np.uint64)
    off[1:] = np.cumsum([len(x) for x in hs])
    return b, off


def write_results(directory, out, top, read_headers, ref_headers, cands=None, taxids=None,
                  first_read_id=0, thresholds=None, append=False, all_groups=False):
    """The reference's four result logs for one batch (mh_write_results).
    Headers as MirAge keeps them (b">name\0")."""
    lib = load()
    out = np.ascontiguousarray(out, dtype=READ_RESULT_DTYPE)
    top = np.ascontiguousarray(top, dtype=np.uint32)
    rh, rho = _headers(read_headers)
    fh, fho = _headers(ref_headers)
    c = None if cands is None else np.ascontiguousarray(cands, dtype=CAND_DTYPE)
    tx = None if taxids is None else np.ascontiguousarray(taxids, dtype=np.int32)
    thr = None if thresholds is None else ctypes.byref(LogThresholds(*thresholds))
    flags = (MH_RESULTS_APPEND if append else 0) | (MH_RESULTS_ALL_GROUPS if all_groups else 0)
    st = lib.mh_write_results(os.fsencode(str(directory)), first_read_id, len(out), _np_ptr(rh),
                              _np_ptr(rho), _np_ptr(out), _np_ptr(top), _np_ptr(c),
                              0 if c is None else len(c), len(ref_headers), _np_ptr(fh),
                              _np_ptr(fho), _np_ptr(tx), thr, flags)
    if st != MH_OK:
        raise MirageError(st, last_error())


def last_error():
    """Error text of the calling thread's last call made without a handle."""
    return load().mh_last_error(None).decode()


class MhmFile:
    """An opened .mhm v4 hashmap file (mh_mhm_open); host-side only."""

    def __init__(self, path):
        self.lib = load()
        m = ctypes.c_void_p()
        st = self.lib.mh_mhm_open(os.fsencode(path), ctypes.byref(m))
        if st != MH_OK:
            raise MirageError(st, last_error())
        self.m = m

    def close(self):
        if self.m:
            self.lib.mh_mhm_close(self.m)
            self.m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self):
        i = MhmInfo()
        st = self.lib.mh_mhm_get_info(self.m, ctypes.byref(i))
        if st != MH_OK:
            raise MirageError(st, last_error())
        return {f: getattr(i, f) for f, _ in MhmInfo._fields_}

    def references(self):
        """(sequence bytes, n_refs+1 offsets, header byte strings, taxonomy ids)."""
        i = self.info()
        n = i["n_refs"]
        rb = np.zeros(max(i["ref_bytes"], 1), np.uint8)
        ro = np.zeros(n + 1, np.uint64)
        hb = np.zeros(max(i["header_bytes"], 1), np.uint8)
        ho = np.zeros(n + 1, np.uint64)
        tx = np.zeros(max(n, 1), np.int32)
        st = self.lib.mh_mhm_references(self.m, _np_ptr(rb), _np_ptr(ro), _np_ptr(hb), _np_ptr(ho),
                                        _np_ptr(tx))
        if st != MH_OK:
            raise MirageError(st, last_error())
        heads = [hb[int(ho[r]):int(ho[r + 1])].tobytes() for r in range(n)]
        return rb[:i["ref_bytes"]], ro, heads, tx[:n]

    def bucket(self, h, last_bucket_empty=True):
        """Entries (ref, pos) of bucket h in file order, or None when absent."""
        n = ctypes.c_int64()
        fl = MH_INDEX_LAST_BUCKET_EMPTY if last_bucket_empty else 0
        st = self.lib.mh_mhm_bucket(self.m, h, fl, None, None, 0, ctypes.byref(n))
        if st != MH_OK:
            raise MirageError(st, last_error())
        if n.value < 0:
            return None
        refs = np.zeros(max(n.value, 1), np.uint32)
        pos = np.zeros(max(n.value, 1), np.uint32)
        st = self.lib.mh_mhm_bucket(self.m, h, fl, _np_ptr(refs), _np_ptr(pos), n.value,
                                    ctypes.byref(n))
        if st != MH_OK:
            raise MirageError(st, last_error())
        return list(zip(refs[:n.value].tolist(), pos[:n.value].tolist()))


class Server:
    """Serving loop over a resident index (mh_server_start)."""

    def __init__(self, index, hbm_budget=0, max_inflight=0, max_batch_reads=0, flags=0):
        self.index = index
        self.lib = index.handle.lib
        cfg = ServerConfig(hbm_budget, max_inflight, max_batch_reads, flags, 0)
        s = ctypes.c_void_p()
        index.handle.check(self.lib.mh_server_start(index.handle.h, index.ix, ctypes.byref(cfg),
                                                    ctypes.byref(s)), "mh_server_start")
        self.s = s
        self.sizes = {}

    def submit(self, reads, read_off):
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        read_off = np.ascontiguousarray(read_off, dtype=np.uint64)
        rid = ctypes.c_uint64()
        st = self.lib.mh_server_submit(self.s, _np_ptr(reads), _np_ptr(read_off),
                                       len(read_off) - 1, ctypes.byref(rid))
        if st != MH_OK:
            raise MirageError(st, "mh_server_submit")
        self.sizes[rid.value] = len(read_off) - 1
        return rid.value

    def submit_arrays(self, reads, read_off):
        """submit() without copies or checks: reads (uint8) and read_off
        (uint64) must be contiguous numpy arrays; the library copies them."""
        rid = ctypes.c_uint64()
        n = len(read_off) - 1
        st = self.lib.mh_server_submit(self.s, reads.ctypes.data, read_off.ctypes.data, n,
                                       ctypes.byref(rid))
        if st != MH_OK:
            raise MirageError(st, "mh_server_submit")
        self.sizes[rid.value] = n
        return rid.value

    def wait_into(self, rid, out, top, timeout=-1.0):
        """wait() into caller buffers (out: READ_RESULT_DTYPE of at least the
        request's reads; top: uint32).  Returns (top ids needed, info)."""
        n = self.sizes.pop(rid)
        need = ctypes.c_uint64()
        info = RequestInfo()
        st = self.lib.mh_server_wait(self.s, rid, timeout, out.ctypes.data, top.ctypes.data,
                                     len(top), ctypes.byref(need), ctypes.byref(info))
        if st != MH_OK:
            raise MirageError(st, f"request {rid}")
        return int(need.value), {"submit_s": info.submit_s, "admit_s": info.admit_s,
                                 "done_s": info.done_s, "n_reads": n}

    def wait(self, rid, timeout=-1.0):
        """Returns (results, top pool, info dict) of request rid."""
        n = self.sizes.pop(rid)
        out = np.zeros(n, READ_RESULT_DTYPE)
        cap = max(64, 8 * n)
        while True:
            top = np.zeros(cap, np.uint32)
            need = ctypes.c_uint64()
            info = RequestInfo()
            st = self.lib.mh_server_wait(self.s, rid, timeout, _np_ptr(out), _np_ptr(top), cap,
                                         ctypes.byref(need), ctypes.byref(info))
            if st == MH_ERR_OVERFLOW:
                cap = int(need.value) + 16
                continue
            if st != MH_OK:
                raise MirageError(st, f"request {rid}")
            return out, top[:need.value], {"submit_s": info.submit_s, "admit_s": info.admit_s,
                                           "done_s": info.done_s, "n_reads": info.n_reads}

    def stop(self):
        if self.s:
            self.lib.mh_server_stop(self.s)
            self.s = None

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass


def tie_sets(out, top):
    """Per-read sorted top-tie tuples from a result array and its pool."""
    res = []
    for r in out:
        n = int(r["n_top"])
        o = int(r["top_off"])
        res.append(tuple(sorted(top[o:o + n].tolist())) if n else ())
    return res
