""".mhm v4 files on the device: decode into HBM, classification through the
loaded index, and the writer.

The fixtures are hashmap files the reference wrote and then classified the
golden reads through; loading them must give the index the library builds
from the same references, bucket for bucket, and the reference's printed
results for every read.  Writing the library's index back must reproduce the
reference's file byte for byte.
"""
import numpy as np
import pytest

from mirage_amd import datagen as D
from mirage_amd import native
from oracle import mhm as M
from tests import _data
from tests.test_gpu_parity import compare_cands


pytestmark = pytest.mark.gpu

NAMES = _data.mhm_names()


def same_results(out, top, bout, btop, label):
    """Two device results bit for bit: score bits, strand, candidate count,
    tie set and status of every read."""
    ties, bties = native.tie_sets(out, top), native.tie_sets(bout, btop)
    bad = [i for i in range(len(out))
           if (np.float32(out["highest"][i]).view(np.uint32) !=
               np.float32(bout["highest"][i]).view(np.uint32) or
               out["strand"][i] != bout["strand"][i] or out["n_cand"][i] != bout["n_cand"][i]
               or out["status"][i] != bout["status"][i] or ties[i] != bties[i])]
    assert not bad, f"{label}: {len(bad)} reads differ, first {bad[0]}"


def info_of(ix):
    i = ix.info()
    i.pop("device_bytes")
    return i


@pytest.mark.parametrize("last_empty", [True, False])
@pytest.mark.parametrize("name", NAMES)
def test_load_equals_build(handle, tmp_path, name, last_empty):
    """The decoded index holds the same buckets as the one built from the
    FASTA references (the T = 30 file has a bucket 0 of 1527 entries, which
    takes the workgroup decode)."""
    fx, rb, ro, _, _ = _data.load_golden(name)
    print(f"load {name} last_empty={last_empty}", flush=True)
    ixf = handle.load_mhm(_data.mhm_file(name, tmp_path), last_bucket_empty=last_empty)
    ixb = handle.build_index(rb, ro, fx["threshold"], last_bucket_empty=last_empty)
    try:
        assert info_of(ixf) == info_of(ixb)
        m = M.read(_data.mhm_bytes(name), last_bucket_empty=last_empty)
        assert ixf.info()["n_entries"] == sum(len(v) for v in m.buckets.values())
        for h, ents in m.buckets.items():
            assert ixf.bucket(h) == sorted(ents), h
        absent = next(h for h in range(1, 1 << 28) if h not in m.buckets)
        assert ixf.bucket(absent) is None
    finally:
        ixf.free()
        ixb.free()


@pytest.mark.parametrize("name", NAMES)
def test_classify_through_loaded_file(handle, tmp_path, name):
    """Reads classified through the reference's own hashmap file give the
    reference's printed lines, and bit-identical results and candidates to
    the index built from the references."""
    fx, rb, ro, qb, qo = _data.load_golden(name)
    print(f"classify {name}", flush=True)
    ixf = handle.load_mhm(_data.mhm_file(name, tmp_path))
    ixb = handle.build_index(rb, ro, fx["threshold"])
    exp = _data.expected_lines(fx, len(qo) - 1)
    try:
        for flags in (0, native.MH_CLASSIFY_BATCHED):
            out, top, cands = ixf.classify_host(qb, qo, flags=flags, want_cands=True)
            ties = native.tie_sets(out, top)
            for i, e in enumerate(exp):
                got = None if out["n_top"][i] == 0 else ("%f" % out["highest"][i], ties[i])
                assert got == e, f"{name} read {i}: loaded {got} reference {e}"
            bout, btop, bcands = ixb.classify_host(qb, qo, flags=flags, want_cands=True)
            same_results(out, top, bout, btop, f"{name} loaded vs built")
            compare_cands(cands, bcands, f"{name} loaded vs built")
    finally:
        ixf.free()
        ixb.free()


@pytest.mark.parametrize("name", NAMES)
def test_save_reproduces_reference_file(handle, tmp_path, name):
    fx, rb, ro, _, _ = _data.load_golden(name)
    print(f"save {name}", flush=True)
    ix = handle.build_index(rb, ro, fx["threshold"], last_bucket_empty=False)
    try:
        path = tmp_path / "saved.mhm"
        ix.save_mhm(str(path), headers=_data.fasta_headers(fx)[1])
        assert path.read_bytes() == _data.mhm_bytes(name)
    finally:
        ix.free()


def test_save_refuses_index_without_last_bucket(handle, tmp_path):
    fx, rb, ro, _, _ = _data.load_golden("t22_tiny_mixed")
    ix = handle.build_index(rb, ro, fx["threshold"], last_bucket_empty=True)
    try:
        with pytest.raises(native.MirageError) as e:
            ix.save_mhm(str(tmp_path / "x.mhm"))
        assert e.value.status == -7
    finally:
        ix.free()


def test_threshold_rules(handle, tmp_path):
    """threshold 0 takes the file's, a smaller one is refused, a larger one
    is kept for the queries."""
    path = _data.mhm_file("t22_tiny_mixed", tmp_path)
    with pytest.raises(native.MirageError):
        handle.load_mhm(path, threshold=20)
    ix = handle.load_mhm(path, threshold=24)
    assert ix.info()["threshold"] == 24
    ix.free()


def test_round_trip_at_size(handle, tmp_path):
    """Workload 2's reference model with 300 references (about 15 Mbp, k = 16),
    8-byte hashes and stream positions: save, load, and classify 20k reads
    through both indexes; the file re-read by the host reader holds the same
    buckets as the device."""
    w = D.workload(2, 0.02)
    w.refs.n = 300
    rb, ro, qb, qo, _ = D.generate(w)
    print(f"round trip: {len(ro) - 1} refs, {len(rb)} bases", flush=True)
    ixb = handle.build_index(rb, ro, w.threshold, last_bucket_empty=False)
    path = str(tmp_path / "w2.mhm")
    ixb.save_mhm(path, hash_bytes=8, streampos_bytes=8)
    ixf = handle.load_mhm(path, last_bucket_empty=False)
    try:
        assert info_of(ixf) == info_of(ixb)
        with native.MhmFile(path) as f:
            i = f.info()
            assert (i["hash_bytes"], i["streampos_bytes"]) == (8, 8)
            rng = np.random.default_rng(7)
            for h in rng.integers(0, 1 << 28, 2000).tolist():
                b = f.bucket(h, last_bucket_empty=False)
                assert (sorted(b) if b is not None else None) == ixb.bucket(h)
        for flags in (0, native.MH_CLASSIFY_BATCHED):
            out, top, cands = ixf.classify_host(qb, qo, flags=flags, want_cands=True)
            bout, btop, bcands = ixb.classify_host(qb, qo, flags=flags, want_cands=True)
            same_results(out, top, bout, btop, "round trip")
            compare_cands(cands, bcands, "round trip")
    finally:
        ixf.free()
        ixb.free()


def test_corrupt_records_are_errors(handle, tmp_path):
    """A reference id past the last reference in a one-thread record, and run
    counts that overrun the entry count in the workgroup-decoded bucket 0,
    fail the load with MH_ERR_ARG; the decode writes only inside each
    bucket's own entries."""
    b = _data.mhm_bytes("t30_tiny_nruns")
    m = M.read(b, last_bucket_empty=False)
    # bucket 0: 1527 (0, 0) entries, one run; its count field follows the ids
    sp0 = m.streampos[0]
    assert m.hashes[0] == 0 and len(m.buckets[0]) > 64
    bpo, bid, bpos, bns = b[sp0:sp0 + 4]
    n = len(m.buckets[0])
    run_at = sp0 + 4 + 2 * bpo + bns + n * bid + 1 + bpos
    cases = []
    x = bytearray(b)
    x[run_at] = (x[run_at] + 1) & 0xFF  # count one more than the bucket holds
    cases.append(("big record", x))
    i = next(j for j in range(1, len(m.hashes) - 1) if b[m.streampos[j]] == 1)
    x = bytearray(b)
    x[m.streampos[i] + 4] = 0xFF  # refIDMin 255 > 6 references
    cases.append(("small record", x))
    for what, data in cases:
        print(f"corrupt {what}", flush=True)
        p = tmp_path / "bad.mhm"
        p.write_bytes(bytes(data))
        with pytest.raises(native.MirageError) as e:
            handle.load_mhm(str(p), last_bucket_empty=False)
        assert e.value.status == -1 and "corrupt record" in str(e.value), what
    ix = handle.load_mhm(_data.mhm_file("t30_tiny_nruns", tmp_path), last_bucket_empty=False)
    assert ix.bucket(0) == [(0, 0)] * len(m.buckets[0])
    ix.free()
