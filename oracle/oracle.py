"""ctypes binding of the CPU oracle (oracle/mirage_oracle.cpp).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py for the CPU-baseline leg.  The product package
(mirage_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmirage_oracle.so")

MO_FLAG_LAST_BUCKET_EMPTY = 1
MO_FLAG_ULTRAFAST = 1


class ReadResult(ctypes.Structure):
    _fields_ = [("highest", ctypes.c_float), ("strand", ctypes.c_uint32),
                ("n_top", ctypes.c_uint32), ("n_cand", ctypes.c_uint32)]


CAND_DTYPE = np.dtype([("read", "<u4"), ("strand", "<u4"), ("ref", "<u4"), ("hits", "<u4"),
                       ("sum_r", "<u8"), ("sum_c", "<u8"), ("score", "<f4"), ("rel", "<f4"),
                       ("n_seeds", "<u4"), ("pad", "<u4")])
RESULT_DTYPE = np.dtype([("highest", "<f4"), ("strand", "<u4"), ("n_top", "<u4"),
                         ("n_cand", "<u4")])
WORK_FIELDS = ("probes", "hits", "entries", "ext_compares", "seeds", "cands")


class Work(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint64) for f in WORK_FIELDS]


def build():
    """Compile the oracle with make (g++ -O3)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.mo_hash32.restype = ctypes.c_uint32
        L.mo_hash32.argtypes = [P, ctypes.c_int32]
        L.mo_hash28.restype = ctypes.c_uint32
        L.mo_hash28.argtypes = [P, ctypes.c_int32]
        L.mo_index_build.restype = P
        L.mo_index_build.argtypes = [P, P, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint32]
        L.mo_index_build_for_reads.restype = P
        L.mo_index_build_for_reads.argtypes = [P, P, ctypes.c_uint32, ctypes.c_int32,
                                               ctypes.c_uint32, P, P, ctypes.c_uint32]
        L.mo_index_free.argtypes = [P]
        L.mo_index_num_entries.restype = ctypes.c_uint64
        L.mo_index_num_entries.argtypes = [P]
        L.mo_index_num_buckets.restype = ctypes.c_uint64
        L.mo_index_num_buckets.argtypes = [P]
        L.mo_index_bucket.restype = ctypes.c_int64
        L.mo_index_bucket.argtypes = [P, ctypes.c_uint32, P, P, ctypes.c_uint64]
        L.mo_index_ref_cutoff.restype = ctypes.c_uint64
        L.mo_index_ref_cutoff.argtypes = [P, ctypes.c_uint32]
        L.mo_classify.restype = ctypes.c_int32
        L.mo_classify.argtypes = [P, P, P, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint32,
                                  ctypes.c_int32, P, P, ctypes.c_uint32, P, ctypes.c_uint64, P]
        L.mo_score_seeds.restype = ctypes.c_float
        L.mo_score_seeds.argtypes = [P, ctypes.c_uint64, P, ctypes.c_uint64, P, ctypes.c_uint64,
                                     ctypes.c_int32, ctypes.c_uint32, P, P, P]
        L.mo_relative_score.restype = ctypes.c_float
        L.mo_relative_score.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.mo_std_sort_desc.restype = None
        L.mo_std_sort_desc.argtypes = [P, P, ctypes.c_uint64]
        L.mo_std_stable_sort_tax.restype = None
        L.mo_std_stable_sort_tax.argtypes = [P, ctypes.c_uint64, P]
        L.mo_count_work.restype = ctypes.c_int32
        L.mo_count_work.argtypes = [P, P, P, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint32,
                                    ctypes.c_int32, ctypes.POINTER(Work)]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def hash28(seq: bytes) -> int:
    b = np.frombuffer(seq, dtype=np.uint8)
    return lib().mo_hash28(_ptr(b), len(b))


def hash32(seq: bytes) -> int:
    b = np.frombuffer(seq, dtype=np.uint8)
    return lib().mo_hash32(_ptr(b), len(b))


def relative_score(qsize, rsize, score):
    return lib().mo_relative_score(qsize, rsize, score)


def score_seeds(query, ref, seeds, threshold, ultrafast=False):
    """seeds: (n,4) uint32 array of (start.ref, start.query, last.ref, last.query)."""
    q = np.ascontiguousarray(np.frombuffer(bytes(query), dtype=np.uint8))
    r = np.ascontiguousarray(np.frombuffer(bytes(ref), dtype=np.uint8))
    s = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint32).reshape(-1, 4))
    sr = ctypes.c_uint64()
    sc = ctypes.c_uint64()
    h = ctypes.c_uint32()
    score = lib().mo_score_seeds(_ptr(q), len(q), _ptr(r), len(r), _ptr(s), len(s), threshold,
                                 MO_FLAG_ULTRAFAST if ultrafast else 0, ctypes.byref(sr),
                                 ctypes.byref(sc), ctypes.byref(h))
    return score, sr.value, sc.value, h.value


class Index:
    """Reference index built by the oracle (Hashmap.cpp build semantics)."""

    def __init__(self, ref_bytes: np.ndarray, ref_off: np.ndarray, threshold: int,
                 last_bucket_empty: bool = True):
        self.ref_bytes = np.ascontiguousarray(ref_bytes, dtype=np.uint8)
        self.ref_off = np.ascontiguousarray(ref_off, dtype=np.uint64)
        self.threshold = threshold
        self.n_refs = len(self.ref_off) - 1
        flags = MO_FLAG_LAST_BUCKET_EMPTY if last_bucket_empty else 0
        self._h = lib().mo_index_build(_ptr(self.ref_bytes), _ptr(self.ref_off), self.n_refs,
                                       threshold, flags)
        if not self._h:
            raise ValueError("oracle index build failed")

    @classmethod
    def for_reads(cls, ref_bytes, ref_off, threshold, reads, read_off,
                  last_bucket_empty: bool = True):
        """The index restricted to the buckets these reads probe
        (mo_index_build_for_reads): the same results for these reads as the
        full index, for references too large to index whole on the host.
        T >= 20 only."""
        self = cls.__new__(cls)
        self.ref_bytes = np.ascontiguousarray(ref_bytes, dtype=np.uint8)
        self.ref_off = np.ascontiguousarray(ref_off, dtype=np.uint64)
        self.threshold = threshold
        self.n_refs = len(self.ref_off) - 1
        r = np.ascontiguousarray(reads, dtype=np.uint8)
        ro = np.ascontiguousarray(read_off, dtype=np.uint64)
        flags = MO_FLAG_LAST_BUCKET_EMPTY if last_bucket_empty else 0
        self._h = lib().mo_index_build_for_reads(_ptr(self.ref_bytes), _ptr(self.ref_off),
                                                 self.n_refs, threshold, flags, _ptr(r), _ptr(ro),
                                                 len(ro) - 1)
        if not self._h:
            raise ValueError("oracle index build for reads failed (T < 20?)")
        return self

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mo_index_free(self._h)
            self._h = None

    @property
    def num_entries(self):
        return lib().mo_index_num_entries(self._h)

    @property
    def num_buckets(self):
        return lib().mo_index_num_buckets(self._h)

    def cutoff(self, ref):
        return lib().mo_index_ref_cutoff(self._h, ref)

    def bucket(self, h):
        n = lib().mo_index_bucket(self._h, h, None, None, 0)
        if n < 0:
            return None
        refs = np.zeros(max(n, 1), np.uint32)
        pos = np.zeros(max(n, 1), np.uint32)
        lib().mo_index_bucket(self._h, h, _ptr(refs), _ptr(pos), n)
        return list(zip(refs[:n].tolist(), pos[:n].tolist()))

    def classify(self, reads: np.ndarray, read_off: np.ndarray, n_threads=0, top_cap=64,
                 want_cands=False, ultrafast=False):
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        read_off = np.ascontiguousarray(read_off, dtype=np.uint64)
        n = len(read_off) - 1
        out = np.zeros(n, RESULT_DTYPE)
        top = np.zeros((n, top_cap), np.uint32)
        ncand = ctypes.c_uint64(0)
        flags = MO_FLAG_ULTRAFAST if ultrafast else 0
        cands = None
        if want_cands:
            rc = lib().mo_classify(self._h, _ptr(reads), _ptr(read_off), n, self.threshold,
                                   flags, n_threads, _ptr(out), _ptr(top), top_cap, None, 0,
                                   ctypes.byref(ncand))
            cands = np.zeros(ncand.value, CAND_DTYPE)
        rc = lib().mo_classify(self._h, _ptr(reads), _ptr(read_off), n, self.threshold, flags,
                               n_threads, _ptr(out), _ptr(top), top_cap, _ptr(cands),
                               0 if cands is None else len(cands), ctypes.byref(ncand))
        if rc != 0:
            raise ValueError(f"oracle classify failed ({rc})")
        return out, top, cands

    def count_work(self, reads, read_off, n_threads=0, ultrafast=False):
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        read_off = np.ascontiguousarray(read_off, dtype=np.uint64)
        w = Work()
        rc = lib().mo_count_work(self._h, _ptr(reads), _ptr(read_off), len(read_off) - 1,
                                 self.threshold, MO_FLAG_ULTRAFAST if ultrafast else 0,
                                 n_threads, ctypes.byref(w))
        if rc != 0:
            raise ValueError("oracle count_work failed")
        return {f: getattr(w, f) for f in WORK_FIELDS}


def std_sort_desc(scores):
    """libstdc++ std::sort of (float32 score, id) pairs, score descending."""
    sc = np.ascontiguousarray([s for s, _ in scores], dtype=np.float32)
    ids = np.ascontiguousarray([i for _, i in scores], dtype=np.uint32)
    lib().mo_std_sort_desc(_ptr(sc), _ptr(ids), len(sc))
    return list(zip(sc.tolist(), ids.tolist()))


def std_stable_sort_tax(ids, taxids):
    """libstdc++ std::stable_sort of reference ids with the result logs'
    taxonomy comparator."""
    v = np.ascontiguousarray(ids, dtype=np.uint64)
    t = np.ascontiguousarray(taxids, dtype=np.int32)
    lib().mo_std_stable_sort_tax(_ptr(v), len(v), _ptr(t))
    return v.tolist()


def top_sets(out, top):
    """Per-read sorted tie sets as tuples.  The top array must have held
    every tie set (classify's top_cap at least the largest n_top)."""
    if len(out) and int(out["n_top"].max()) > top.shape[1]:
        raise ValueError(f"a tie set of {int(out['n_top'].max())} references exceeds top_cap "
                         f"{top.shape[1]}: classify again with a larger top_cap")
    return [tuple(top[i, :out["n_top"][i]].tolist()) for i in range(len(out))]
