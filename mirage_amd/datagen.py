"""Seeded synthetic workloads (references and reads) for tests and the benchmark.

Base seed 0x4D495241 + workload number; xoshiro256** streams per reference
and per read (mirage_amd/csrc/mh_datagen.cpp), so data does not depend on the
thread count.  k-mer length k maps to the reference's threshold T = 2(k-1)
(stride T/2, k = T - T/2 + 1, Analyser.cpp:1387-1389).
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from . import _build

BASE_SEED = 0x4D495241


def threshold_for_k(k: int) -> int:
    return 2 * (k - 1)


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = _build.datagen_lib_path()
        if not os.path.exists(path):
            _build.build_datagen()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        u64, dbl = ctypes.c_uint64, ctypes.c_double
        L.mhg_lengths.argtypes = [u64, ctypes.c_int, dbl, dbl, dbl, dbl, dbl, u64, u64, u64, P]
        L.mhg_refs.argtypes = [u64, P, dbl, dbl, u64, u64, u64, dbl, u64, P]
        L.mhg_reads.argtypes = [P, P, u64, u64, P, dbl, dbl, dbl, dbl, dbl, u64, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class LengthSpec:
    mode: int = 0          # 0 fixed, 1 lognormal, 2 uniform, 3 uniform/lognormal mixture
    a: float = 150
    b: float = 0
    c: float = 0
    d: float = 0
    mix: float = 0
    lo: int = 1
    cap: int = 1 << 40

    def draw(self, n, seed):
        out = np.zeros(n, np.uint64)
        lib().mhg_lengths(n, self.mode, self.a, self.b, self.c, self.d, self.mix, self.lo,
                          self.cap, seed, _p(out))
        return out


@dataclass
class RefSpec:
    n: int
    lengths: LengthSpec
    gc_lo: float = 0.5
    gc_hi: float = 0.5
    seg_pool: int = 0
    seg_lo: int = 1000
    seg_hi: int = 20000
    seg_frac: float = 0.0
    total: int = 0         # when > 0, lengths are rescaled to this total


@dataclass
class ReadSpec:
    n: int
    lengths: LengthSpec
    p_exact: float = 0.0
    p_random: float = 0.0
    sub_rate: float = 0.0
    indel_rate: float = 0.0
    rc_frac: float = 0.0


@dataclass
class Workload:
    name: str
    k: int
    refs: RefSpec
    reads: ReadSpec
    seed: int = BASE_SEED
    notes: dict = field(default_factory=dict)

    @property
    def threshold(self):
        return threshold_for_k(self.k)


def offsets(lengths):
    off = np.zeros(len(lengths) + 1, np.uint64)
    np.cumsum(lengths, out=off[1:])
    return off


def make_refs(spec: RefSpec, seed: int):
    lens = spec.lengths.draw(spec.n, seed)
    if spec.total:
        scale = spec.total / float(lens.sum())
        lens = np.maximum(np.round(lens.astype(np.float64) * scale), 1000).astype(np.uint64)
    off = offsets(lens)
    by = np.empty(int(off[-1]), np.uint8)
    lib().mhg_refs(spec.n, _p(off), spec.gc_lo, spec.gc_hi, spec.seg_pool, spec.seg_lo,
                   spec.seg_hi, spec.seg_frac, seed, _p(by))
    return by, off


def make_reads(spec: ReadSpec, ref_bytes, ref_off, seed: int):
    lens = spec.lengths.draw(spec.n, seed + 0x5EED)
    off = offsets(lens)
    by = np.empty(int(off[-1]), np.uint8)
    truth = np.zeros((spec.n, 3), np.uint32)
    lib().mhg_reads(_p(ref_bytes), _p(ref_off), len(ref_off) - 1, spec.n, _p(off), spec.p_exact,
                    spec.p_random, spec.sub_rate, spec.indel_rate, spec.rc_frac, seed + 0x5EED,
                    _p(by), _p(truth))
    return by, off, truth


def generate(w: Workload):
    rb, ro = make_refs(w.refs, w.seed)
    qb, qo, truth = make_reads(w.reads, rb, ro, w.seed)
    return rb, ro, qb, qo, truth


def _ref_config2():
    return RefSpec(n=5000, lengths=LengthSpec(mode=1, a=50000, b=1.0, lo=1000, cap=2_000_000),
                   gc_lo=0.40, gc_hi=0.60)


def workload(num: int, scale: float = 1.0) -> Workload:
    """Workloads 1-5.  scale < 1 shrinks read counts (and index size for 4, 5)
    for tests; 1.0 is the full size."""
    s = BASE_SEED + num
    if num == 1:
        return Workload("plain-small: 100 refs x 10 kb, 10k x 150 bp, k=12", 12,
                        RefSpec(n=100, lengths=LengthSpec(mode=0, a=10000)),
                        ReadSpec(n=max(1, int(10000 * scale)), lengths=LengthSpec(mode=0, a=150),
                                 p_random=0.10, sub_rate=0.01), seed=s)
    if num == 2:
        return Workload("plain-gpu: 5000 refs lognormal(50 kb, 1.0), 1M x 150 bp, k=16", 16,
                        _ref_config2(),
                        ReadSpec(n=max(1, int(1_000_000 * scale)),
                                 lengths=LengthSpec(mode=0, a=150), p_exact=0.05, p_random=0.10,
                                 sub_rate=0.02), seed=s)
    if num == 3:
        return Workload("batched-mixed: 5000 refs, 2M reads 75 bp-20 kb, k=16", 16,
                        _ref_config2(),
                        ReadSpec(n=max(1, int(2_000_000 * scale)),
                                 lengths=LengthSpec(mode=3, a=75, b=300, c=2000, d=1.0, mix=0.8,
                                                    lo=75, cap=20000),
                                 sub_rate=0.03, indel_rate=0.005), seed=BASE_SEED + 2,
                        notes={"long_tail_sigma": 1.0, "index_seed": "workload 2"})
    if num == 4:
        return Workload("served: 20000 refs ~2 Gbp, reads 100-5000 bp, k=16", 16,
                        RefSpec(n=max(10, int(20000 * scale)),
                                lengths=LengthSpec(mode=1, a=50000, b=1.0, lo=1000,
                                                   cap=2_000_000),
                                total=int(2_000_000_000 * scale)),
                        ReadSpec(n=max(1, int(100_000 * scale)),
                                 lengths=LengthSpec(mode=2, a=100, b=5000), sub_rate=0.02,
                                 indel_rate=0.005), seed=s)
    if num == 5:
        return Workload("large-repetitive: 200000 refs ~8 Gbp, 10M x 150 bp, k=20", 20,
                        RefSpec(n=max(10, int(200_000 * scale)),
                                lengths=LengthSpec(mode=1, a=40000, b=0.5, lo=2000,
                                                   cap=2_000_000),
                                total=int(8_000_000_000 * scale), seg_pool=1000,
                                seg_lo=1000, seg_hi=20000, seg_frac=0.30),
                        ReadSpec(n=max(1, int(10_000_000 * scale)),
                                 lengths=LengthSpec(mode=0, a=150), sub_rate=0.01), seed=s)
    raise ValueError(num)
