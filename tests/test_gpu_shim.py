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
