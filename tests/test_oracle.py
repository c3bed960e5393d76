"""CPU tests of the oracle: known answers, reference goldens, the reference's
own functions (through oracle/_ref/ref_harness when it is built), a second
transcription, and mutation sensitivity of the goldens."""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests import _data
from tests import _model

HARNESS = os.path.join(O.HERE, "_ref", "ref_harness")


def fbits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def test_fnv1_known_answers():
    # FNV-1 32-bit test vectors (multiply, then xor).
    assert O.hash32(b"") == 0x811C9DC5
    assert O.hash32(b"a") == 0x050C5D7E
    assert O.hash32(b"foobar") == 0x31F0B262


def test_fold_to_28_bits():
    for s in (b"", b"a", b"foobar", b"ACGTACGTACGTACGT", b"N" * 12):
        h = O.hash32(s)
        assert O.hash28(s) == ((h >> 28) ^ h) & 0x0FFFFFFF
    assert O.hash28(b"N" * 12) == 0x0995D49E


@pytest.mark.parametrize("name", _data.golden_names())
def test_oracle_matches_reference_binary(name):
    fx, rb, ro, qb, qo = _data.load_golden(name)
    ix = O.Index(rb, ro, fx["threshold"])
    out, top, _ = ix.classify(qb, qo)
    exp = _data.expected_lines(fx, len(qo) - 1)
    ties = O.top_sets(out, top)
    for i, e in enumerate(exp):
        got = None if out["n_top"][i] == 0 else ("%f" % out["highest"][i], ties[i])
        assert got == e, f"{name} read {i}: oracle {got} reference {e}"


def test_goldens_cover_both_outcomes():
    total = classified = 0
    for name in _data.golden_names():
        fx, _, _, _, qo = _data.load_golden(name)
        total += len(qo) - 1
        classified += len(fx["expected"])
        ties = sum(1 for v in fx["expected"].values() if len(v[1]) > 1)
        if name == "t38_repetitive":
            assert ties > 0  # shared segments produce tied references
    assert 0 < classified < total


def test_single_exact_seed_closed_form():
    # One seed of length L at the start of the query whose tail meets a
    # mismatch at once: R = L(L+1)/2, C = L, H = L.
    L, T = 40, 30
    q = b"A" * L + b"C" * 10
    ref = b"G" + b"A" * L + b"T" * 10
    score, sr, sc, h = O.score_seeds(q, ref, [[1, 0, 1 + L, L]], T)
    assert (sr, sc, h) == (L * (L + 1) // 2, L, L)
    sf = np.float32(np.float32(sr) / np.float32(sc))
    pct = np.float32((100.0 / float(np.float32(len(q)))) * float(np.float32(h)))
    assert fbits(score) == fbits(float(np.float32((float(sf) / 100.0) * float(pct))))


def test_head_compares_off_diagonal():
    # Query row rs-1 is compared with reference byte start.Reference (one
    # base off the seed diagonal, MatrixCPU.cpp:127-132).
    T = 10
    q = b"X" * 5 + b"A" * 12
    ref = b"Y" * 4 + b"XA" + b"A" * 12
    # seed: query [5, 17) on reference [5, 17); diagonal offset 0
    _, sr, sc, h = O.score_seeds(q, ref, [[5, 5, 17, 17]], T)
    # q[4] = 'X' vs ref[5] = 'A' -> mismatch, head 0 (on the diagonal ref[4] = 'X' would match)
    assert sc == 12


def test_no_seed_reaching_threshold_scores_minus_two():
    score, sr, sc, h = O.score_seeds(b"ACGT" * 10, b"TTTT" * 10, [[0, 0, 5, 5]], 30)
    assert score == -2.0 and sc == 0


def test_index_invariants():
    rng = np.random.default_rng(3)
    refs = [bytes(rng.choice(list(b"ACGT"), n).tolist()) for n in (50, 300, 1000, 17)]
    refs[2] = refs[2][:600] + b"N" * 20 + refs[2][620:]
    rb, ro = _data.pack([r.decode() for r in refs])
    T = 22
    k = T - T // 2 + 1
    ix = O.Index(rb, ro, T, last_bucket_empty=False)
    cut = [ix.cutoff(r) for r in range(len(refs))]
    assert cut[0] == 50 - k and cut[1] == 300 - k and cut[3] == 17 - k
    assert cut[2] == 600 - k + 1 or cut[2] <= 600  # first all-N window ends the reference
    # windows after the cut of reference 2 come back as (0, 0) entries of
    # bucket 0 (the reference converts its reserved size, Hashmap.cpp:1252)
    phantom = (1000 - k) - cut[2]
    assert phantom > 0
    assert ix.num_entries == sum(cut) + phantom
    assert (ix.bucket(0) or []).count((0, 0)) == phantom
    # brute force: every indexed window is in its bucket
    for r, seq in enumerate(refs):
        for i in range(0, cut[r], 37):
            assert (r, i) in ix.bucket(O.hash28(seq[i:i + k]))


def test_fragment_rule_below_t20():
    rng = np.random.default_rng(5)
    seq = bytes(rng.choice(list(b"ACGT"), 4000).tolist())
    rep = b"ACGTTGCA" * 200
    rb, ro = _data.pack([seq.decode(), rep.decode(), rep.decode()])
    m18 = _model.Model([seq, rep, rep], 18, last_bucket_empty=False)
    ix = O.Index(rb, ro, 18, last_bucket_empty=False)
    assert ix.num_buckets == len(m18.buckets)
    assert ix.num_entries == sum(len(v) for v in m18.buckets.values())


def test_model_agrees_with_oracle_on_golden():
    fx, rb, ro, qb, qo = _data.load_golden("t22_tiny_mixed")
    m = _model.Model([r.encode() for r in fx["refs"]], fx["threshold"])
    exp = _data.expected_lines(fx, len(qo) - 1)
    for i in range(0, len(qo) - 1, 5):
        assert m.classify(fx["reads"][i].encode()) == exp[i], i


@pytest.mark.parametrize("mutation", _model.MUTATIONS)
def test_goldens_reject_planted_mistakes(mutation):
    """Every planted mistake changes at least one golden line."""
    for name in ("t30_engineered", "t22_tiny_mixed", "t30_rc_copies", "t22_low_complexity", "t30_variant_copies",
                 "t38_repetitive", "t30_n_runs", "t22_uniform"):
        fx, rb, ro, qb, qo = _data.load_golden(name)
        m = _model.Model([r.encode() for r in fx["refs"]], fx["threshold"], mutation)
        exp = _data.expected_lines(fx, len(qo) - 1)
        for i in range(len(qo) - 1):
            if m.classify(fx["reads"][i].encode()) != exp[i]:
                return
    pytest.fail(f"mutation {mutation} not detected by the goldens")


def _harness(lines):
    p = subprocess.run([HARNESS], input="\n".join(lines) + "\n", capture_output=True, text=True,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"), timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    return p.stdout.split("\n")


@pytest.mark.skipif(not os.path.exists(HARNESS), reason="oracle/_ref not built (make -C oracle ref)")
def test_scoring_matches_reference_calculate_score():
    rng = np.random.default_rng(11)
    cases, lines = [], []
    for _ in range(400):
        T = int(rng.choice([18, 22, 30, 38]))
        n = int(rng.integers(60, 400))
        ref = rng.choice(list(b"ACGT"), n).astype(np.uint8)
        m = int(rng.integers(T + 5, 250))
        q = rng.choice(list(b"ACGT"), m).astype(np.uint8)
        seeds = []
        for _ in range(int(rng.integers(1, 5))):
            L = int(rng.integers(1, min(m, n) // 2))
            qs = int(rng.integers(0, m - L))
            rs = int(rng.integers(0, n - L))
            q[qs:qs + L] = ref[rs:rs + L]
            if rng.random() < 0.5 and qs > 0 and rs > 1:  # make the off-diagonal head match
                q[qs - 1] = ref[rs]
            seeds.append((rs, qs, rs + L, qs + L))
        seeds.sort(key=lambda t: (t[1], t[0]))
        uf = int(rng.random() < 0.1)
        cases.append((bytes(q), bytes(ref), seeds, T, uf))
        lines.append(f"S {T} {uf} {m} {n} {len(seeds)}")
        lines.append(bytes(q).decode())
        lines.append(bytes(ref).decode())
        lines.append(" ".join(str(v) for sd in seeds for v in sd))
    out = _harness(lines)
    for (q, ref, seeds, T, uf), line in zip(cases, out):
        score, _, _, _ = O.score_seeds(q, ref, seeds, T, ultrafast=bool(uf))
        assert f"{fbits(score):08x}" == line.split()[0]


@pytest.mark.skipif(not os.path.exists(HARNESS), reason="oracle/_ref not built (make -C oracle ref)")
def test_relative_score_and_hash_match_reference():
    rng = np.random.default_rng(13)
    cases, lines = [], []
    for _ in range(500):
        qs = float(rng.integers(20, 20000))
        rs = float(rng.integers(20, 2_000_000))
        sc = float(np.float32(rng.random() * qs / 2))
        cases.append((qs, rs, sc))
        lines.append(f"R {int(qs)} {int(rs)} {fbits(sc):08x}")
    seqs = [bytes(rng.choice(list(b"ACGTN"), int(rng.integers(1, 40))).tolist()) for _ in range(200)]
    for s in seqs:
        lines.append(f"H {len(s)} {s.decode()}")
    out = _harness(lines)
    for (qs, rs, sc), line in zip(cases, out[:500]):
        assert f"{fbits(O.relative_score(qs, rs, sc)):08x}" == line.strip()
    for s, line in zip(seqs, out[500:]):
        assert O.hash28(s) == int(line)


def _same_results(a, b):
    return (np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
            and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]))


@pytest.mark.parametrize("name", [n for n in _data.golden_names()
                                  if int(n.split("_")[0][1:]) >= 20])
def test_index_for_reads_gives_full_index_results(name):
    """The index restricted to the probed buckets (used where the full
    index does not fit host memory) classifies the reads exactly as the
    full index does, per read and per candidate, on every golden with
    T >= 20 (N runs, bucket 0's copies, shared segments, odd T)."""
    fx, rb, ro, qb, qo = _data.load_golden(name)
    full = O.Index(rb, ro, fx["threshold"])
    part = O.Index.for_reads(rb, ro, fx["threshold"], qb, qo)
    assert part.num_entries < full.num_entries
    a = full.classify(qb, qo, want_cands=True)
    b = part.classify(qb, qo, want_cands=True)
    assert _same_results(a, b)


def test_index_for_reads_buckets_and_last_bucket():
    """Bucket for bucket equal to the full index on every probed hash (and
    absent elsewhere); the highest hash of the whole index is present but
    empty when a read probes it (MO_FLAG_LAST_BUCKET_EMPTY); bucket 0's
    dropped-window copies survive; T < 20 is refused."""
    rng = np.random.default_rng(21)
    refs = [bytes(rng.choice(list(b"ACGT"), n).tolist()) for n in (3000, 2500, 4000)]
    refs[1] = refs[1][:900] + b"N" * 30 + refs[1][930:]  # a cutoff: bucket 0 copies
    T = 30
    s, k = T // 2, T - T // 2 + 1
    rb, ro = _data.pack([r.decode() for r in refs])
    full = O.Index(rb, ro, T)
    # the window with the highest hash of the whole index, found by brute force
    best = max(((O.hash28(seq[i:i + k]), r, i) for r, seq in enumerate(refs)
                for i in range(full.cutoff(r))), key=lambda t: t[0])
    top_read = refs[best[1]][best[2]:best[2] + k] + b"ACGT" * 10
    reads = [top_read] + [refs[r][a:a + 150] for r, a in ((0, 100), (1, 700), (2, 2000))]
    reads.append(b"A" * k + b"C" * 40)  # probes hashes the references lack
    qb, qo = _data.pack([x.decode() for x in reads])
    part = O.Index.for_reads(rb, ro, T, qb, qo)
    assert full.bucket(best[0]) == [] and part.bucket(best[0]) == []
    comp = bytes.maketrans(b"ACGT", b"TGCA")
    probed = set()
    for x in reads:
        for strand in (x, x.translate(comp)[::-1]):
            for j in range(0, len(strand) - T, s):
                probed.add(O.hash28(strand[j:j + k]))
    for h in probed:
        assert part.bucket(h) == full.bucket(h), h
    unprobed = [O.hash28(refs[2][i:i + k]) for i in range(3000, 3050)]
    assert all(part.bucket(h) is None for h in unprobed if h not in probed)
    assert _same_results(full.classify(qb, qo, want_cands=True),
                         part.classify(qb, qo, want_cands=True))
    with pytest.raises(ValueError):
        O.Index.for_reads(rb, ro, 18, qb, qo)
