"""Oracle parity at scale: the large repetitive index (k = 20, shared
segments, 2 Gbp), the mixed-length batch on workload 2's full 5,000-reference
index, and the served request stream of workload 4 (0.5 Gbp), each compared
read by read and candidate by candidate with the oracle (MatrixCPU.cpp:323-370,
Analyser.cpp:592-611).  The oracle indexes only the buckets the reads probe
(Index.for_reads, equal to the full index for them), so references of
gigabases stay within host memory.
"""
import numpy as np
import pytest

from mirage_amd import datagen as D
from mirage_amd import native
from oracle import oracle as O
from tests.test_gpu_parity import compare_cands, compare_reads
from tests.test_gpu_tiers import mutate, pack, shared_segment

pytestmark = pytest.mark.gpu

TOP = 1024  # oracle tie-set room: segments shared by ~100 references


@pytest.fixture(scope="module")
def repetitive():
    """Workload 5's recipe at a quarter of its index: 50,000 references,
    2 Gbp, 30 % with segments from the 1,000-segment pool; 20,000 reads."""
    w = D.workload(5, 0.25)
    rb, ro = D.make_refs(w.refs, w.seed)
    w.reads.n = 20_000
    qb, qo, _ = D.make_reads(w.reads, rb, ro, w.seed)
    oix = O.Index.for_reads(rb, ro, w.threshold, qb, qo)
    oout, otop, ocands = oix.classify(qb, qo, want_cands=True, top_cap=TOP)
    del oix
    return w, rb, ro, qb, qo, (oout, otop, ocands)


@pytest.mark.parametrize("mode", ["defaults", "hbm_tiers", "plain", "preprobe",
                                  "preprobe_hbm_tiers"])
def test_repetitive_index_matches_oracle(handle, repetitive, mode):
    """Every read and candidate equals the oracle's, with the index-sized
    default tiles, with LDS tiers too small for most strands (so they finish
    in the HBM-workspace tiers), and in the plain form; with the probe pass,
    whose records the overflow tiers read for the strands the first tile
    passes on."""
    w, rb, ro, qb, qo, (oout, otop, ocands) = repetitive
    assert ro[-1] >= 2_000_000_000 * 0.99
    ix = handle.build_index(rb, ro, w.threshold)
    info = ix.info()
    epp = info["n_entries"] / info["n_buckets"]
    assert epp > 7
    flags = 0 if mode == "plain" else native.MH_CLASSIFY_BATCHED
    if mode.startswith("preprobe"):
        flags |= native.MH_CLASSIFY_PREPROBE
    try:
        if mode.endswith("hbm_tiers"):
            handle.set_tiles([(200, 16, 32, 0), (handle.max_read_len, 64, 64, 0)])
            handle.set_overflow_tiles([(256, 64, 32, 0), (handle.max_read_len, 64, 32, 0)])
        out, top, cands = ix.classify_host(qb, qo, flags=flags, want_cands=True)
        rows = handle.kernel_counters()
        nout, ntop, _ = ix.classify_host(qb, qo, flags=flags)
    finally:
        handle.set_tiles(None)
        handle.set_overflow_tiles(None)
        ix.free()
    compare_reads(out, top, oout, otop, f"repetitive {mode}")
    compare_reads(nout, ntop, oout, otop, f"repetitive {mode}, no candidate output")
    compare_cands(cands, ocands, f"repetitive {mode}")
    assert max(len(t) for t in native.tie_sets(out, top)) > 16  # segment ties
    if mode.endswith("hbm_tiers"):
        assert rows[10]["queries"] > 0.5 * 2 * (len(qo) - 1)  # most strands in HBM workspaces
    if mode == "preprobe":
        assert rows[8]["queries"] > 0.05 * 2 * (len(qo) - 1)  # the second-chance tier reads records


def test_mixed_lengths_on_the_full_index(handle):
    """The mixed-length batch (75 bp - 20 kb, 3 % substitutions, 0.5 %
    indels) against workload 2's full 5,000-reference index, batched and
    plain, read by read and candidate by candidate."""
    w2 = D.workload(2, 1.0)
    rb, ro = D.make_refs(w2.refs, w2.seed)
    w3 = D.workload(3)
    w3.reads.n = 1500
    qb, qo, _ = D.make_reads(w3.reads, rb, ro, w3.seed)
    assert np.diff(qo).max() > 8000
    oout, otop, ocands = O.Index.for_reads(rb, ro, w3.threshold, qb, qo).classify(
        qb, qo, want_cands=True, top_cap=TOP)
    ix = handle.build_index(rb, ro, w3.threshold)
    for flags in (native.MH_CLASSIFY_BATCHED, 0):
        out, top, cands = ix.classify_host(qb, qo, flags=flags, want_cands=True)
        compare_reads(out, top, oout, otop, f"mixed lengths, flags {flags}")
        compare_cands(cands, ocands, f"mixed lengths, flags {flags}")
    ix.free()


def _serve(ix, reqs, **cfg):
    srv = native.Server(ix, **cfg)
    ids = [srv.submit(b, o) for b, o in reqs]
    res = [srv.wait(i) for i in ids]
    srv.stop()
    return res


def _check_served(res, reqs, oout, otop, label):
    """Each request's results against the oracle's rows for its reads."""
    at = 0
    for (out, top, info), (b, o) in zip(res, reqs):
        n = len(o) - 1
        assert info["n_reads"] == n
        compare_reads(out, top, oout[at:at + n], otop[at:at + n], f"{label} request @{at}")
        at += n


@pytest.mark.parametrize("flags", [0, native.MH_CLASSIFY_PREPROBE])
def test_served_requests_match_oracle(handle, flags):
    """Workload 4's recipe at 0.5 Gbp (5,000 references, k = 16): 60
    requests of log-uniform size (1-4,000 reads of 100-5,000 bp) served
    together through three slots, without and with the probe pass; every
    request's every read equals the oracle's."""
    w = D.workload(4, 0.25)
    rb, ro = D.make_refs(w.refs, w.seed)
    assert ro[-1] >= 0.49e9
    rng = np.random.default_rng(7)
    sizes = np.floor(10 ** rng.uniform(0, np.log10(4000), 60)).astype(int)
    pool = D.ReadSpec(n=int(sizes.sum()), lengths=D.LengthSpec(mode=2, a=100, b=5000),
                      sub_rate=0.02, indel_rate=0.005)
    qb, qo, _ = D.make_reads(pool, rb, ro, w.seed)
    reqs, at = [], 0
    for n in sizes:
        reqs.append((qb[int(qo[at]):int(qo[at + n])], qo[at:at + n + 1] - qo[at]))
        at += n
    oout, otop, _ = O.Index.for_reads(rb, ro, w.threshold, qb, qo).classify(qb, qo, top_cap=TOP)
    h = native.Handle(0, 8192)
    ix = h.build_index(rb, ro, w.threshold)
    res = _serve(ix, reqs, hbm_budget=4 << 30, max_inflight=3, max_batch_reads=8192, flags=flags)
    _check_served(res, reqs, oout, otop, f"served, flags {flags}")
    ix.free()
    h.close()


def test_served_batches_requeued_when_tie_pools_overflow(handle):
    """Batches whose tied references outgrow the slot's pools (16 reads per
    batch, 300-way ties: the top-tie pool of 8 x 16 + 1,024 ids overflows)
    are requeued with grown pools; the results still equal the oracle's."""
    rb, ro, seg = shared_segment(300, 2400, 500, seed=31)
    rng = np.random.default_rng(32)
    reads = []
    for i in range(96):
        if i % 3:
            a = int(rng.integers(0, len(seg) - 150))
            reads.append(mutate(rng, seg[a:a + 150], 1))
        else:
            r = int(rng.integers(0, 300))
            a = int(ro[r]) + int(rng.integers(0, 350))
            reads.append(rb[a:a + 150].copy())
    qb, qo = pack(reads)
    reqs = [(qb[int(qo[a]):int(qo[a + 8])], qo[a:a + 9] - qo[a]) for a in range(0, 96, 8)]
    oout, otop, _ = O.Index(rb, ro, 30).classify(qb, qo, top_cap=TOP)
    assert max(oout["n_top"]) == 300
    ix = handle.build_index(rb, ro, 30)
    res = _serve(ix, reqs, hbm_budget=1 << 30, max_inflight=2, max_batch_reads=16)
    _check_served(res, reqs, oout, otop, "requeued")
    ix.free()
