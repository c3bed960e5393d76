"""The C++ call surface (matrix_hip.hpp) on the GPU: Classify and the
scoring-only GetHighestScore against the oracle."""
import os
import struct
import subprocess

import numpy as np
import pytest

from mirage_amd import _build
from oracle import oracle as O
from tests import _data

pytestmark = pytest.mark.gpu


def fbits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def test_shim_classify_and_score_seeds():
    fx, rb, ro, qb, qo = _data.load_golden("t30_gc_lognormal")
    T = fx["threshold"]
    refs, reads = fx["refs"], fx["reads"][:200]
    rng = np.random.default_rng(3)
    queries = []
    for _ in range(120):
        q = rng.choice(list(b"ACGT"), int(rng.integers(60, 300))).astype(np.uint8)
        cands = []
        for _ in range(int(rng.integers(1, 4))):
            r = int(rng.integers(0, len(refs)))
            ref = np.frombuffer(refs[r].encode(), np.uint8)
            seeds = []
            for _ in range(int(rng.integers(1, 4))):
                L = int(rng.integers(5, min(len(q), len(ref)) // 2))
                a = int(rng.integers(0, len(q) - L))
                b = int(rng.integers(0, len(ref) - L))
                q[a:a + L] = ref[b:b + L]
                seeds.append((b, a, b + L, a + L))
            seeds.sort(key=lambda t: t[1])
            cands.append((r, seeds))
        queries.append((bytes(q).decode(), cands))
    lines = [f"{T} {len(refs)} {len(reads)} {len(queries)}"] + refs + reads
    for q, cands in queries:
        lines.append(q)
        lines.append(str(len(cands)))
        for r, seeds in cands:
            lines.append(f"{r} {len(seeds)}")
            lines.append(" ".join(str(v) for sd in seeds for v in sd))
    p = subprocess.run([_build.shim_driver_path()], input="\n".join(lines) + "\n",
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    out = p.stdout.strip().split("\n")
    c_lines = [ln for ln in out if ln.startswith("C ")]
    s_lines = [ln for ln in out if ln.startswith("S ")]
    # Classify vs oracle
    rbb, roo = _data.pack(refs)
    qbb, qoo = _data.pack(reads)
    oout, otop, _ = O.Index(rbb, roo, T).classify(qbb, qoo)
    ties = O.top_sets(oout, otop)
    for i, ln in enumerate(c_lines):
        f = ln.split()
        assert int(f[1], 16) == fbits(oout["highest"][i]), i
        assert tuple(sorted(int(x) for x in f[3:])) == ties[i], i
    # GetHighestScore vs oracle scoring + CreateResultsScores
    for (q, cands), ln in zip(queries, s_lines):
        best, scores = -1.0, []
        for r, seeds in cands:
            sc = O.score_seeds(q.encode(), refs[r].encode(), seeds, T)[0]
            if sc > -1:
                rel = O.relative_score(float(len(q)), float(len(refs[r])), sc)
                best = rel if rel > best else best
                scores.append((fbits(rel), r))
        f = ln.split()
        assert int(f[1], 16) == fbits(best)
        got = [(int(f[3 + 2 * k], 16), int(f[4 + 2 * k])) for k in range(int(f[2]))]
        assert got == scores


def test_score_seeds_checks_every_seed(handle):
    """mh_score_seeds refuses malformed pairs before scoring anything
    (MH_ERR_ARG, or MH_ERR_RANGE for a query longer than the handle's
    maximum), naming the first bad pair; well-formed seeds score as the
    oracle does, and seeds with start.query > last.query are skipped
    (MatrixCPU.cpp:57-58)."""
    from mirage_amd import native
    rng = np.random.default_rng(11)
    refs = [bytes(rng.choice(list(b"ACGT"), n).astype(np.uint8)) for n in (400, 300)]
    rb, ro = _data.pack([r.decode() for r in refs])
    T = 30
    ix = handle.build_index(rb, ro, T)
    q = bytearray(rng.choice(list(b"ACGT"), 120).astype(np.uint8))
    q[10:70] = refs[0][100:160]
    queries = np.frombuffer(bytes(q), np.uint8)
    qoff = np.array([0, len(q)], np.uint64)
    good = [(100, 10, 160, 70)]
    skipped = [(390, 50, 300, 40)]  # start.query > last.query: skipped, positions never read

    def call(pair_ref, seeds_per_pair, q_=queries, qo_=qoff):
        so = np.zeros(len(pair_ref) + 1, np.uint64)
        so[1:] = np.cumsum([len(x) for x in seeds_per_pair])
        flat = np.array([v for x in seeds_per_pair for sd in x for v in sd], np.uint32)
        return ix.score_seeds(q_, qo_, np.zeros(len(pair_ref), np.uint32), pair_ref, so, flat)

    score, rel, sr, sc, hits = call([0, 0], [good, good + skipped])
    exp = O.score_seeds(bytes(q), refs[0], good, T)
    for p in range(2):
        assert np.float32(score[p]).view(np.uint32) == np.float32(exp[0]).view(np.uint32)
        assert (int(sr[p]), int(sc[p]), int(hits[p])) == tuple(exp[1:])
    bad_cases = [
        ([0, 2], [good, good], "pair 1", "reference id"),
        ([0, 0, 1], [good, good, [(100, 10, 160, 121)]], "pair 2", "ends past its query"),
        ([1], [[(300, 10, 310, 20)]], "pair 0", "outside its reference"),
        ([1], [[(290, 10, 301, 20)]], "pair 0", "outside its reference"),
    ]
    for pair_ref, seeds, where, what in bad_cases:
        with pytest.raises(native.MirageError) as e:
            call(pair_ref, seeds)
        assert e.value.status == native.MH_ERR_ARG
        assert where in str(e.value) and what in str(e.value), str(e.value)
    long_q = np.frombuffer(bytes(rng.choice(list(b"ACGT"), handle.max_read_len + 1)
                                 .astype(np.uint8)), np.uint8)
    with pytest.raises(native.MirageError) as e:
        call([0], [[]], long_q, np.array([0, len(long_q)], np.uint64))
    assert e.value.status == native.MH_ERR_RANGE and "pair 0" in str(e.value)
    ix.free()
