"""mh_classify_noalloc (the entry point that allocates nothing) against
mh_classify on the same inputs: two calls in a row, the second with fewer
reads, in the plain and the batched form, with strands that reach every
overflow tier (the HBM tiers run on the fixed array's workspaces); and its
refusals (too many reads, graph capture)."""
import numpy as np
import pytest

from mirage_amd import datagen as D
from mirage_amd import native
from tests.test_gpu_tiers import mutate, pack, shared_segment

pytestmark = pytest.mark.gpu


def _results(d_out, d_top, n):
    out = d_out.cpu().numpy().view(native.READ_RESULT_DTYPE)[:n].copy()
    top = d_top.cpu().numpy().view(np.uint32)
    return out, native.tie_sets(out, top)


def _equal(a, b):
    (oa, ta), (ob, tb) = a, b
    return (np.array_equal(oa["highest"].view(np.uint32), ob["highest"].view(np.uint32))
            and np.array_equal(oa["strand"], ob["strand"])
            and np.array_equal(oa["n_cand"], ob["n_cand"])
            and np.array_equal(oa["status"], ob["status"]) and ta == tb)


def _inputs():
    """Workload-1 reads plus reads of a 2.4 kb segment shared by 300
    references (20k-47k entries per strand: HBM tiers, 300-way ties)."""
    w = D.workload(1, 0.05)
    rb1, ro1, qb1, qo1, _ = D.generate(w)
    rb, ro, seg = shared_segment(300, 2400, 500, seed=11)
    rng = np.random.default_rng(12)
    reads = [qb1[int(qo1[i]):int(qo1[i + 1])] for i in range(len(qo1) - 1)]
    for ln in (1000, 1500, 2000, 2400):
        st = int(rng.integers(0, len(seg) - ln + 1))
        reads.append(seg[st:st + ln])
        reads.append(mutate(rng, seg[st:st + ln], 6))
    order = rng.permutation(len(reads))
    qb, qo = pack([reads[i] for i in order])
    # references: the segment set followed by workload 1's
    refs = np.concatenate([rb, rb1])
    off = np.concatenate([ro, ro1[1:] + ro[-1]]).astype(np.uint64)
    return refs, off, qb, qo, w.threshold


@pytest.mark.parametrize("flags", [0, native.MH_CLASSIFY_BATCHED])
def test_noalloc_matches_classify(handle, flags):
    import torch
    rb, ro, qb, qo, T = _inputs()
    ix = handle.build_index(rb, ro, T)
    n = len(qo) - 1
    d_reads = torch.from_numpy(qb).cuda()
    d_off = torch.from_numpy(qo.view(np.int64)).cuda()
    cap = 400 * n + 1024
    for nn in (n, n // 3):  # a second call with fewer reads
        got = []
        for noalloc in (False, True):
            d_out = torch.full((nn * native.READ_RESULT_DTYPE.itemsize,), 0xAB,
                               dtype=torch.uint8, device="cuda")
            d_top = torch.zeros(cap * 4, dtype=torch.uint8, device="cuda")
            call = ix.classify_noalloc if noalloc else ix.classify_device
            tn, _ = call(d_reads, d_off, nn, d_out, d_top, cap, flags=flags)
            assert tn <= cap
            got.append(_results(d_out, d_top, nn))
            if noalloc:
                c = handle.counters()
        assert _equal(got[0], got[1]), f"{nn} reads, flags {flags}"
        if nn == n:
            assert c["overflow"] > 0  # strands went through the overflow tiers
            assert max(len(t) for t in got[1][1]) == 300
    ix.free()


def test_noalloc_refusals(handle):
    import torch
    w = D.workload(1, 0.01)
    rb, ro, qb, qo, _ = D.generate(w)
    ix = handle.build_index(rb, ro, w.threshold)
    d_reads = torch.from_numpy(qb).cuda()
    d_off = torch.from_numpy(qo.view(np.int64)).cuda()
    d_out = torch.zeros(64 * native.READ_RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    d_top = torch.zeros(1024, dtype=torch.uint8, device="cuda")
    with pytest.raises(native.MirageError) as e:
        ix.classify_noalloc(d_reads, d_off, native.MH_NOALLOC_MAX_READS + 1, d_out, d_top, 256)
    assert e.value.status == native.MH_ERR_RANGE and "at most" in str(e.value)
    with pytest.raises(native.MirageError) as e:
        ix.classify_noalloc(d_reads, d_off, 10, d_out, d_top, 256,
                            flags=native.MH_CLASSIFY_BATCHED | native.MH_CLASSIFY_GRAPH)
    assert e.value.status == native.MH_ERR_ARG
    ix.free()
