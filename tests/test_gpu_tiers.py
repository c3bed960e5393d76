"""Overflow tiers and per-read statuses of the HIP path.

A strand whose working set exceeds its length tile moves to a middle LDS
tier, then to HBM workspaces (a moderate one over many workgroups, then the
largest); the result must not depend on which tier ran.  Reads the path
cannot classify are flagged in their status word instead of failing the
batch.
"""
import numpy as np
import pytest

from mirage_amd import datagen as D
from mirage_amd import native
from oracle import oracle as O
from tests.test_gpu_parity import compare_cands, compare_reads

pytestmark = pytest.mark.gpu

ACGT = np.frombuffer(b"ACGT", np.uint8)


def shared_segment(n_refs, seg_len, flank, seed):
    """n_refs references, each a shared segment between random flanks."""
    rng = np.random.default_rng(seed)
    seg = rng.choice(ACGT, seg_len)
    parts, off = [], [0]
    for _ in range(n_refs):
        r = np.concatenate([rng.choice(ACGT, flank), seg, rng.choice(ACGT, flank)])
        parts.append(r)
        off.append(off[-1] + len(r))
    return np.concatenate(parts), np.array(off, np.uint64), seg


def pack(reads):
    off = np.zeros(len(reads) + 1, np.uint64)
    off[1:] = np.cumsum([len(r) for r in reads])
    return np.concatenate(reads).astype(np.uint8), off


def mutate(rng, r, n):
    r = r.copy()
    for p in rng.integers(0, len(r), n):
        r[p] = ACGT[(np.searchsorted(ACGT, r[p]) + 1) % 4]
    return r


@pytest.mark.parametrize("flags", [0, native.MH_CLASSIFY_BATCHED])
def test_repeat_reads_reach_every_tier(handle, flags):
    """300 references share a 2.4 kb segment: reads from it gather 20k-47k
    bucket entries per strand, beyond the LDS tiers and, for the longest,
    beyond the first HBM tier (2^15 entries); 300-way ties fill the tie pool."""
    rb, ro, seg = shared_segment(300, 2400, 500, seed=11)
    rng = np.random.default_rng(12)
    reads = []
    for ln in (1000, 1200, 1500, 1800, 2000, 2400):
        st = int(rng.integers(0, len(seg) - ln + 1))
        reads.append(seg[st:st + ln])
        reads.append(mutate(rng, seg[st:st + ln], 6))
    for i in range(6):  # flank reads with a single reference
        a = int(ro[i * 40]) + int(rng.integers(0, 300))
        reads.append(rb[a:a + 150])
    qb, qo = pack(reads)
    ix = handle.build_index(rb, ro, 30)
    out, top, cands = ix.classify_host(qb, qo, flags=flags, want_cands=True)
    c = handle.counters()
    assert c["entries"] > 3 * (1 << 15)
    assert c["overflow"] >= 12
    oix = O.Index(rb, ro, 30)
    oout, otop, ocands = oix.classify(qb, qo, want_cands=True, top_cap=512)
    compare_reads(out, top, oout, otop, "shared segment")
    compare_cands(cands, ocands, "shared segment")
    assert max(len(t) for t in native.tie_sets(out, top)) == 300


TILE_TABLES = {
    "block64": [(320, 64, 64, 0), (None, 64, 64, 0)],
    "lanes": [(200, 16, 32, 0), (600, 32, 32, 0), (None, 64, 64, 0)],
    "lanes8": [(300, 8, 32, 0), (600, 32, 32, 0), (None, 64, 64, 0)],
}


@pytest.mark.parametrize("flags", [0, native.MH_CLASSIFY_BATCHED])
@pytest.mark.parametrize("table", sorted(TILE_TABLES))
def test_tiny_tiles_send_work_to_the_tiers(handle, flags, table):
    """With small tiles (whole-workgroup teams, or lane teams of 8, 16 and
    32 lanes) many strands overflow; results are unchanged."""
    try:
        handle.set_tiles([(ml or handle.max_read_len, b, e, g)
                          for ml, b, e, g in TILE_TABLES[table]])
        w = D.workload(3, 0.0005)
        w.refs.n = 300
        rb, ro, qb, qo, _ = D.generate(w)
        ix = handle.build_index(rb, ro, w.threshold)
        out, top, cands = ix.classify_host(qb, qo, flags=flags, want_cands=True)
        assert handle.counters()["overflow"] > 0
    finally:
        handle.set_tiles(None)
    oix = O.Index(rb, ro, w.threshold)
    oout, otop, ocands = oix.classify(qb, qo, want_cands=True)
    compare_reads(out, top, oout, otop, "tiny tiles")
    compare_cands(cands, ocands, "tiny tiles")


def test_read_statuses():
    """Bad bases, reads over the handle's length limit and strands with more
    entries than the largest workspace (2^19; here 1598 probes x 340
    references) are flagged per read; the other reads of the batch are
    classified as usual."""
    h = native.Handle(0, 25000)
    rb, ro, seg = shared_segment(340, 24000, 300, seed=21)
    rng = np.random.default_rng(22)
    good = [seg[100:400].copy(), rb[int(ro[5]) + 10:int(ro[5]) + 260].copy()]
    bad = seg[500:800].copy()
    bad[150] = ord("X")
    lower = np.frombuffer(bytes(seg[1000:1300]).lower(), np.uint8).copy()
    reads = [good[0], bad, rng.choice(ACGT, 25001), seg[:24000].copy(), good[1], lower]
    qb, qo = pack(reads)
    ix = h.build_index(rb, ro, 30)
    for flags in (0, native.MH_CLASSIFY_BATCHED):
        out, top, _ = ix.classify_host(qb, qo, flags=flags)
        st = [int(x) for x in out["status"]]
        assert st == [0, native.MH_READ_BADBASE, native.MH_READ_TOO_LONG,
                      native.MH_READ_CAPACITY, 0, 0], st
        for i in (1, 2):
            assert out["n_top"][i] == 0 and out["n_cand"][i] == 0
        keep = [0, 4, 5]
        sb, so = pack([reads[i] for i in keep])
        oout, otop = O.Index(rb, ro, 30).classify(sb, so, top_cap=512)[:2]
        sub = out[keep]
        ties = native.tie_sets(out, top)
        oties = O.top_sets(oout, otop)
        for j, i in enumerate(keep):
            assert np.float32(sub["highest"][j]).view(np.uint32) == \
                np.float32(oout["highest"][j]).view(np.uint32)
            assert ties[i] == oties[j]
    h.close()


@pytest.mark.parametrize("shift", [1, 2, 3])
def test_preprobe_reads_at_unaligned_addresses(handle, shift):
    """The probe pass loads reads as aligned dwords: a read buffer that starts
    `shift` bytes past an aligned address and ends inside a dword gives the
    same results as the oracle."""
    import torch
    w = D.workload(1, 0.05)
    rb, ro, qb, qo, _ = D.generate(w)
    if len(qb) % 4 == 0:  # make the buffer end inside a dword
        qb, qo = qb[:-1].copy(), qo.copy()
        qo[-1] -= 1
    n = len(qo) - 1
    ix = handle.build_index(rb, ro, w.threshold)
    host = torch.zeros(len(qb) + 8, dtype=torch.uint8)
    host[shift:shift + len(qb)] = torch.from_numpy(qb)
    d_all = host.cuda()
    d_reads = d_all[shift:shift + len(qb)]  # a view: data pointer + shift
    assert d_reads.data_ptr() % 4 == shift
    d_off = torch.from_numpy(qo.view(np.int64)).cuda()
    d_out = torch.empty(n * native.READ_RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    top_cap = 64 * n + 1024
    d_top = torch.empty(top_cap * 4, dtype=torch.uint8, device="cuda")
    ix.classify_device(d_reads, d_off, n, d_out, d_top, top_cap,
                       flags=native.MH_CLASSIFY_BATCHED | native.MH_CLASSIFY_PREPROBE)
    out = d_out.cpu().numpy().view(native.READ_RESULT_DTYPE)
    top = d_top.cpu().numpy().view(np.uint32)
    oout, otop, _ = O.Index(rb, ro, w.threshold).classify(qb, qo, top_cap=512)
    compare_reads(out, top, oout, otop, f"unaligned by {shift}")


def test_invalid_tile_tables_are_refused(handle):
    """A tile table the kernels cannot run is refused with its reason, and
    the handle keeps its tiles."""
    w = D.workload(1, 0.01)
    rb, ro, _, _, _ = D.generate(w)
    ix = handle.build_index(rb, ro, w.threshold)
    saved = handle.get_tiles(ix)
    bad = [[(handle.max_read_len, 16, 16, 0)],                      # entries_cap below 32
           [(handle.max_read_len, 24, 32, 0)],                      # block not a power of two
           [(400, 16, 32, 0), (200, 32, 64, 0)],                    # max_len not ascending
           [(200, 16, 32, 0)]]                                      # last below max_read_len
    for t in bad:
        with pytest.raises(native.MirageError) as e:
            handle.set_tiles(t)
        assert e.value.status == native.MH_ERR_ARG and "mh_set_tiles" in str(e.value), t
    assert handle.get_tiles(ix) == saved
    ix.free()


def test_default_tiers_set_back_and_refusals(handle):
    """The default overflow tiers that mh_get_tiles reports (the second-chance
    capacity fitted to residency, a multiple of 16) are accepted back by
    mh_set_overflow_tiles with the same results as the oracle; capacities that
    are not multiples of 16 are refused."""
    w = D.workload(1, 0.05)
    rb, ro, qb, qo, _ = D.generate(w)
    ix = handle.build_index(rb, ro, w.threshold)
    oout, otop, _ = O.Index(rb, ro, w.threshold).classify(qb, qo, top_cap=512)
    try:
        _, tiers = handle.get_tiles(ix)
        assert all(t[2] % 16 == 0 and t[2] >= 32 for t in tiers), tiers
        handle.set_overflow_tiles([tuple(t) for t in tiers])
        out, top, _ = ix.classify_host(qb, qo, flags=native.MH_CLASSIFY_BATCHED)
        compare_reads(out, top, oout, otop, "default tiers set back")
        with pytest.raises(native.MirageError) as e:
            handle.set_overflow_tiles([(1024, 256, 840, 0), (handle.max_read_len, 256, 4096, 0)])
        assert e.value.status == native.MH_ERR_ARG and "mh_set_overflow_tiles" in str(e.value)
    finally:
        handle.set_overflow_tiles(None)
        ix.free()


def test_calls_of_2_30_reads_are_refused(handle):
    """Strand ids (2 * read + strand) and the overflow flag share 32 bits: a
    call of 2^30 reads is refused before the device is touched."""
    import torch
    w = D.workload(1, 0.01)
    rb, ro, _, _, _ = D.generate(w)
    ix = handle.build_index(rb, ro, w.threshold)
    one = torch.zeros(64, dtype=torch.uint8, device="cuda")
    with pytest.raises(native.MirageError) as e:
        ix.classify_device(one, one, 1 << 30, one, one, 1)
    assert e.value.status == native.MH_ERR_ARG and "per call" in str(e.value)
    ix.free()


WORKGROUP_TABLES = {
    "wg256": [(200, 256, 1024, 0), (None, 256, 1024, 0)],
    "wg128": [(200, 128, 512, 0), (None, 128, 512, 0)],
}


@pytest.mark.parametrize("flags", [0, native.MH_CLASSIFY_BATCHED])
@pytest.mark.parametrize("table", sorted(WORKGROUP_TABLES))
def test_workgroup_teams_score_many_references(handle, flags, table):
    """Short reads in whole-workgroup teams: reads from a segment shared by 40
    references tie across them, so a strand scores its candidates in several
    batches of 16 reference slots, after the hit-mask filter; reads from
    elsewhere have one candidate.  Every result and candidate equals the
    oracle's."""
    rng = np.random.default_rng(21)
    seg = rng.choice(ACGT, 3000)
    parts, off = [], [0]
    for i in range(60):
        r = rng.choice(ACGT, 20000)
        if i % 3 != 2:  # 40 of the 60 references hold the segment
            at = int(rng.integers(0, 20000 - 3000))
            r[at:at + 3000] = seg
        parts.append(r)
        off.append(off[-1] + len(r))
    rb, ro = np.concatenate(parts), np.array(off, np.uint64)
    reads = []
    for i in range(240):
        if i % 2 == 0:
            a = int(rng.integers(0, 3000 - 150))
            reads.append(mutate(rng, seg[a:a + 150].copy(), 1))
        else:
            r = int(rng.integers(0, 60))
            a = int(ro[r]) + int(rng.integers(0, 20000 - 150))
            reads.append(mutate(rng, rb[a:a + 150].copy(), 2))
    qb, qo = pack(reads)
    try:
        handle.set_tiles([(ml or handle.max_read_len, b, e, g)
                          for ml, b, e, g in WORKGROUP_TABLES[table]])
        ix = handle.build_index(rb, ro, 30)
        out, top, cands = ix.classify_host(qb, qo, flags=flags, want_cands=True, top_cap=64 * 240)
        # without the candidate output the relative scores are looked up
        # after the last round of slots
        nout, ntop, _ = ix.classify_host(qb, qo, flags=flags, top_cap=64 * 240)
        ix.free()
    finally:
        handle.set_tiles(None)
    oix = O.Index(rb, ro, 30)
    oout, otop, ocands = oix.classify(qb, qo, want_cands=True, top_cap=64)
    compare_reads(out, top, oout, otop, f"workgroup teams {table}")
    compare_reads(nout, ntop, oout, otop, f"workgroup teams {table}, no candidate output")
    compare_cands(cands, ocands, f"workgroup teams {table}")
    per_strand = np.unique(cands[["read", "strand"]], return_counts=True)[1]
    assert per_strand.max() > 32  # three or more batches of slots


@pytest.mark.parametrize("flags", [0, native.MH_CLASSIFY_BATCHED])
def test_kernel_counters_add_up(handle, flags):
    """Every strand finishes in exactly one kernel: the per-kernel rows'
    strands add up to both strands of every read, and their work fields to
    the call's totals; strands reach the HBM tiers here."""
    rb, ro, seg = shared_segment(300, 2400, 500, seed=11)
    rng = np.random.default_rng(13)
    reads = [seg[a:a + ln] for a, ln in ((0, 1500), (300, 2000), (100, 150), (900, 150))]
    for i in range(40):
        a = int(ro[i * 7]) + int(rng.integers(0, 300))
        reads.append(rb[a:a + int(rng.integers(100, 400))])
    qb, qo = pack(reads)
    ix = handle.build_index(rb, ro, 30)
    ix.classify_host(qb, qo, flags=flags)
    tot = handle.counters()
    rows = handle.kernel_counters()
    assert len(rows) == 12
    assert sum(r["queries"] for r in rows) == 2 * (len(qo) - 1)
    for f in ("probes", "hits", "entries", "seeds", "cands", "direct", "ext_compares"):
        assert sum(r[f] for r in rows) == tot[f], f
    assert rows[10]["overflow"] > 0 and rows[10]["queries"] + rows[11]["queries"] > 0
    ix.free()
