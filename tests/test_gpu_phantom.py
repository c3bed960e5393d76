"""Bucket 0 and the windows dropped after an "N"-window cutoff.

The reference keeps one (ref 0, pos 0) entry of hash 0 per window it drops
after a reference's first all-"N" window (the build converts its reserved
size, MirAge/Hashmap.cpp:1233-1239, :1252).  The library stores one such
entry and counts the others; a read whose probe hashes to 0 must then score
exactly as against the oracle, which stores every copy: each extra copy is a
one-pair seed (0, j, s, j + s) of reference 0.

ACGATGTGGTACCTCT is a 16-mer whose 28-bit FNV-1 hash is 0 (found by search;
the test checks it), so reads holding it at a probe position hit bucket 0.
"""
import numpy as np
import pytest

from mirage_amd import native
from oracle import mhm as M
from oracle import oracle as O
from tests.test_gpu_parity import compare_cands, compare_reads

pytestmark = pytest.mark.gpu

T, S, K = 30, 15, 16
Z = b"ACGATGTGGTACCTCT"
ACGT = np.frombuffer(b"ACGT", np.uint8)


def rand(rng, n):
    return rng.choice(ACGT, n)


def dataset(ref0_starts_with_z, seed):
    """Six references of 2 kb; references 2 and 4 hold a 40-base N run (their
    later windows become (0, 0) entries of bucket 0).  Reference 0 starts with
    Z or not.  Reads hit bucket 0 at probe positions in several ways."""
    rng = np.random.default_rng(seed)
    refs = [rand(rng, 2000) for _ in range(6)]
    if ref0_starts_with_z:
        refs[0][:K] = np.frombuffer(Z, np.uint8)
    for r, at in ((2, 700), (4, 1200)):
        refs[r][at:at + 40] = ord("N")
    z = np.frombuffer(Z, np.uint8)
    reads = []
    for j in (0, S, 2 * S, 4 * S):
        # j bases of noise, then reference 0 from position 0 when it starts
        # with Z (so the copies' seeds run along the read), else Z then ref 0
        body = (refs[0][:150 - j] if ref0_starts_with_z
                else np.concatenate([z, refs[0][K:150 - j]]))
        reads.append(np.concatenate([rand(rng, j), body]))
    for j in (S, 3 * S):  # Z inside a read of reference 1
        r = refs[1][300:450].copy()
        r[j:j + K] = z
        reads.append(r)
    r = refs[0][100:400].copy()  # Z twice, in a read of reference 0
    r[2 * S:2 * S + K] = z
    r[6 * S:6 * S + K] = z
    reads.append(r)
    for i in range(6):  # plain reads, also across the N runs
        a = int(rng.integers(0, 1800))
        reads.append(refs[i][a:a + 150].copy())
    rb = np.concatenate(refs)
    ro = np.zeros(len(refs) + 1, np.uint64)
    ro[1:] = np.cumsum([len(x) for x in refs])
    qb = np.concatenate(reads).astype(np.uint8)
    qo = np.zeros(len(reads) + 1, np.uint64)
    qo[1:] = np.cumsum([len(x) for x in reads])
    return rb, ro, qb, qo


def test_kmer_hashes_to_zero():
    assert O.hash28(Z) == 0
    assert O.hash28(b"N" * K) != 0


@pytest.mark.parametrize("starts_with_z", [True, False])
def test_reads_hitting_bucket_zero(handle, starts_with_z):
    rb, ro, qb, qo = dataset(starts_with_z, seed=31 + starts_with_z)
    oix = O.Index(rb, ro, T)
    copies = oix.bucket(0).count((0, 0))
    assert copies > 1000  # the windows after both N runs
    ix = handle.build_index(rb, ro, T)
    assert ix.info()["n_entries"] == oix.num_entries
    assert ix.bucket(0) == oix.bucket(0)
    oout, otop, ocands = oix.classify(qb, qo, want_cands=True)
    work = oix.count_work(qb, qo)
    # the copies' seeds count: reference 0 scores with them on some reads
    assert any(c["ref"] == 0 and c["n_seeds"] >= copies for c in ocands)
    for flags in (0, native.MH_CLASSIFY_BATCHED,
                  native.MH_CLASSIFY_BATCHED | native.MH_CLASSIFY_PREPROBE,
                  native.MH_CLASSIFY_NO_DIRECT):
        out, top, cands = ix.classify_host(qb, qo, flags=flags, want_cands=True)
        compare_reads(out, top, oout, otop, f"bucket 0, flags {flags}")
        compare_cands(cands, ocands, f"bucket 0, flags {flags}")
        c = handle.counters()
        for f in ("probes", "hits", "entries", "seeds", "cands", "ext_compares"):
            assert c[f] == work[f], (flags, f, c[f], work[f])
    ix.free()


def test_bucket_zero_through_mhm(handle, tmp_path):
    """Saved and loaded again, bucket 0 keeps every copy: the file holds them
    all (as the reference writes them) and the loaded index classifies as the
    built one."""
    rb, ro, qb, qo = dataset(True, seed=33)
    oix = O.Index(rb, ro, T, last_bucket_empty=False)
    ix = handle.build_index(rb, ro, T, last_bucket_empty=False)
    path = str(tmp_path / "b0.mhm")
    ix.save_mhm(path)
    m = M.read(open(path, "rb").read(), last_bucket_empty=False)
    assert sorted(m.buckets[0]) == oix.bucket(0)
    assert len(m.buckets) == oix.num_buckets
    lx = handle.load_mhm(path, last_bucket_empty=False)
    assert lx.info()["n_entries"] == oix.num_entries
    assert lx.bucket(0) == oix.bucket(0)
    a, ta, ca = ix.classify_host(qb, qo, want_cands=True)
    b, tb, cb = lx.classify_host(qb, qo, want_cands=True)
    assert np.array_equal(a["highest"].view(np.uint32), b["highest"].view(np.uint32))
    assert native.tie_sets(a, ta) == native.tie_sets(b, tb)
    compare_cands(ca, cb, "built vs loaded")
    lx.free()
    ix.free()
