""".mhm v4 hashmap files on the host: the oracle's restatement pinned against
files the reference wrote, and the library's host reader against the oracle.

The fixtures are the reference's own hashmap databases for four golden
datasets (tests/golden/make_ref_goldens.py): the reference wrote them with
SaveHashmapTableV3 and classified the golden reads through them.  Their
buckets must equal the oracle's index built from the same references, which
pins the file decoder independently of its own formulas; writing the decoded
file back must give the same bytes, which pins the writer.
"""
import struct

import numpy as np
import pytest

from mirage_amd import native
from oracle import mhm as M
from oracle import oracle as O
from tests import _data

NAMES = _data.mhm_names()


def oracle_index(name, last_bucket_empty):
    fx, rb, ro, _, _ = _data.load_golden(name)
    return fx, O.Index(rb, ro, fx["threshold"], last_bucket_empty=last_bucket_empty)


def test_fixtures_present():
    assert NAMES == ["t18_tiny_mixed", "t22_fasta_forms", "t22_tiny_mixed", "t30_tiny_nruns"]


@pytest.mark.parametrize("name", NAMES)
def test_reference_file_equals_oracle_index(name):
    """Every bucket the reference wrote equals the oracle's bucket, and no
    bucket is missing: the fragment rule (T = 18), the N-window cutoff and
    the (0, 0) entries of bucket 0 it leaves (T = 30 with N runs) included."""
    b = _data.mhm_bytes(name)
    m = M.read(b, last_bucket_empty=False)
    fx, ix = oracle_index(name, False)
    assert m.threshold == fx["threshold"]
    assert m.refs == [r.encode() for r in fx["refs"]]
    assert m.headers == _data.fasta_headers(fx)[1]
    assert m.taxids == [-1] * len(fx["refs"])
    assert len(m.buckets) == ix.num_buckets
    assert sum(len(v) for v in m.buckets.values()) == ix.num_entries
    for h, ents in m.buckets.items():
        assert sorted(ents) == ix.bucket(h), h
        assert [p for _, p in ents] == sorted(p for _, p in ents)  # written by position
    if name == "t30_tiny_nruns":
        assert m.buckets[0].count((0, 0)) > 0


@pytest.mark.parametrize("name", NAMES)
def test_oracle_writer_reproduces_reference_bytes(name):
    b = _data.mhm_bytes(name)
    assert M.write(M.read(b, last_bucket_empty=False)) == b


@pytest.mark.parametrize("name", NAMES)
def test_loader_leaves_last_bucket_empty(name):
    """As loaded, the highest-hash bucket is present and empty, which is the
    oracle's MO_FLAG_LAST_BUCKET_EMPTY index."""
    m = M.read(_data.mhm_bytes(name), last_bucket_empty=True)
    _, ix = oracle_index(name, True)
    top = max(m.buckets)
    assert m.buckets[top] == [] and ix.bucket(top) == []
    assert ix.num_buckets == len(m.buckets)


def test_planted_decoder_mistake_is_caught():
    """A decoder that forgets refPOSMin fails the bucket comparison."""
    b = _data.mhm_bytes("t22_tiny_mixed")
    good = M.read(b, last_bucket_empty=False)
    orig = M.decode_record

    def no_pos_min(buf, o, end):
        ents, o2 = orig(buf, o, end)
        pos_min = min(p for _, p in ents)
        return [(r, p - pos_min) for r, p in ents], o2
    M.decode_record = no_pos_min
    try:
        bad = M.read(b, last_bucket_empty=False)
    finally:
        M.decode_record = orig
    _, ix = oracle_index("t22_tiny_mixed", False)
    diff = sum(sorted(v) != ix.bucket(h) for h, v in bad.buckets.items())
    assert diff > 0.5 * len(bad.buckets)
    assert all(sorted(v) == ix.bucket(h) for h, v in good.buckets.items())


# ---- the library's host reader (no GPU call) --------------------------------

@pytest.mark.parametrize("name", NAMES)
def test_host_reader_matches_oracle(name, tmp_path):
    b = _data.mhm_bytes(name)
    m = M.read(b, last_bucket_empty=False)
    with native.MhmFile(_data.mhm_file(name, tmp_path)) as f:
        info = f.info()
        assert info["n_buckets"] == len(m.hashes)
        assert info["file_bytes"] == len(b)
        assert info["record_bytes"] == len(b) - m.streampos[0]
        assert info["n_refs"] == len(m.refs)
        assert info["threshold"] == m.threshold
        assert (info["hash_bytes"], info["streampos_bytes"]) == (m.hash_bytes, m.sp_bytes)
        rb, ro, heads, tax = f.references()
        assert rb.tobytes() == b"".join(m.refs)
        assert [int(x) for x in np.diff(ro)] == [len(r) for r in m.refs]
        assert heads == m.headers and tax.tolist() == m.taxids
        for h in m.hashes:
            assert f.bucket(h, last_bucket_empty=False) == m.buckets[h], h
        top = m.hashes[-1]
        assert f.bucket(top, last_bucket_empty=True) == []
        absent = next(h for h in range(1, 1 << 28) if h not in m.buckets)
        assert f.bucket(absent) is None


def _record_at(b, name, pick):
    """(hash, record offset) of the first bucket whose decoded entries
    satisfy pick(entries, record bytes)."""
    m = M.read(b, last_bucket_empty=False)
    for i, h in enumerate(m.hashes[:-1]):
        sp, end = m.streampos[i], m.streampos[i + 1]
        if pick(m.buckets[h], b[sp:end]):
            return h, sp
    raise AssertionError("no such bucket")


def _open_fails(tmp_path, data, what):
    p = tmp_path / "bad.mhm"
    p.write_bytes(bytes(data))
    with pytest.raises(native.MirageError) as e:
        native.MhmFile(str(p))
    assert e.value.status == -1 and what in str(e.value), str(e.value)


def test_host_reader_rejects_bad_files(tmp_path):
    b = bytearray(_data.mhm_bytes("t22_tiny_mixed"))
    _open_fails(tmp_path, b[:len(b) // 2], "past the end")
    _open_fails(tmp_path, b[:100], "")
    x = bytearray(b)
    x[0] = ord("X")
    _open_fails(tmp_path, x, "indicator")
    x = bytearray(b)
    struct.pack_into("<i", x, 73 + 8, 5)
    _open_fails(tmp_path, x, "version")
    x = bytearray(b)
    struct.pack_into("<i", x, 73 + 12, 2)
    _open_fails(tmp_path, x, "hash width")
    x = bytearray(b)  # swap the first two TOC hashes
    t0 = 73 + 24
    h0, h1 = struct.unpack_from("<I", x, t0)[0], struct.unpack_from("<I", x, t0 + 8)[0]
    struct.pack_into("<I", x, t0, h1)
    struct.pack_into("<I", x, t0 + 8, h0)
    _open_fails(tmp_path, x, "ascending")


def test_host_reader_rejects_bad_records(tmp_path):
    """Records that would decode to a width of 0, a reference id past the
    last reference, a position past its reference's end, or run counts that
    do not add up to the entry count are errors of mh_mhm_bucket, not
    out-of-range reads; other buckets still decode."""
    b = _data.mhm_bytes("t22_tiny_mixed")
    m = M.read(b, last_bucket_empty=False)
    cases = []
    h, sp = _record_at(b, "t22", lambda e, r: len(e) >= 2)
    x = bytearray(b)
    x[sp + 1] = 0  # id width 0
    cases.append((h, x))
    h, sp = _record_at(b, "t22", lambda e, r: r[0] == 1)
    x = bytearray(b)
    x[sp + 4] = 0xFF  # refIDMin 255 > 8 references
    cases.append((h, x))
    h, sp = _record_at(b, "t22", lambda e, r: r[0] == 2)
    x = bytearray(b)
    x[sp + 6:sp + 8] = b"\xff\xff"  # refPOSMin 65535, past every reference
    cases.append((h, x))
    h, sp = _record_at(b, "t22", lambda e, r: len(e) == 1 and r[:4] == b"\x01\x01\x01\x01")
    x = bytearray(b)
    # 4 widths, refIDMin, refPOSMin, n, one id, bRep, runs, count, pos
    assert x[sp + 7 + 1 + 1 + 1] == 1
    x[sp + 7 + 1 + 1 + 1] = 2  # one run of two entries for a bucket of one
    cases.append((h, x))
    other = next(k for k in m.hashes[:-1] if all(k != c[0] for c in cases))
    for h, x in cases:
        p = tmp_path / "bad.mhm"
        p.write_bytes(bytes(x))
        with native.MhmFile(str(p)) as f:
            with pytest.raises(native.MirageError) as e:
                f.bucket(h, last_bucket_empty=False)
            assert "corrupt record" in str(e.value)
            assert f.bucket(other, last_bucket_empty=False) == m.buckets[other]
