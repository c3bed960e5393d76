"""The hipGraph-captured batched form: every replay equals the eager call,
with eager work and synchronisation between replays; a changed argument
triggers a new capture."""
import numpy as np
import pytest

from mirage_amd import datagen as D
from mirage_amd import native

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager(handle):
    import torch
    w = D.workload(3, 0.001)
    w.refs.n = 200
    rb, ro, qb, qo, _ = D.generate(w)
    ix = handle.build_index(rb, ro, w.threshold)
    n = len(qo) - 1
    d_reads = torch.from_numpy(qb).cuda()
    d_off = torch.from_numpy(qo.view(np.int64)).cuda()
    cap = 16 * n + 64

    def run(flags, nn=n):
        d_out = torch.zeros(nn * native.READ_RESULT_DTYPE.itemsize, dtype=torch.uint8,
                            device="cuda")
        d_top = torch.zeros(cap * 4, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        ix.classify_device(d_reads, d_off, nn, d_out, d_top, cap, flags=flags)
        out = d_out.cpu().numpy().view(native.READ_RESULT_DTYPE)
        top = d_top.cpu().numpy().view(np.uint32)
        return out["highest"].view(np.uint32).copy(), native.tie_sets(out, top)

    eager = run(native.MH_CLASSIFY_BATCHED)
    d_out = torch.zeros(n * native.READ_RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    d_top = torch.zeros(cap * 4, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    for rep in range(4):
        d_out.fill_(0xAB)  # eager work between replays
        torch.cuda.synchronize()
        ix.classify_device(d_reads, d_off, n, d_out, d_top, cap,
                           flags=native.MH_CLASSIFY_BATCHED | native.MH_CLASSIFY_GRAPH)
        out = d_out.cpu().numpy().view(native.READ_RESULT_DTYPE)
        top = d_top.cpu().numpy().view(np.uint32)
        assert np.array_equal(out["highest"].view(np.uint32), eager[0]), rep
        assert native.tie_sets(out, top) == eager[1], rep
        # an eager call on the same handle between replays
        assert run(0)[0].tolist() == eager[0].tolist()
    # fewer reads: captured again
    half = n // 2
    g = run(native.MH_CLASSIFY_BATCHED | native.MH_CLASSIFY_GRAPH, half)
    e = run(native.MH_CLASSIFY_BATCHED, half)
    assert np.array_equal(g[0], e[0]) and g[1] == e[1]


def test_graph_captured_again_after_arena_grows():
    """An eager call or a FASTA parse that grows the handle's scratch arena
    frees the arena a captured call points into; the next graph call
    captures again, and every graph call equals the eager one."""
    import torch
    h = native.Handle(0, 32768)
    w = D.workload(3, 0.001)
    w.refs.n = 200
    rb, ro, qb, qo, _ = D.generate(w)
    ix = h.build_index(rb, ro, w.threshold)
    n = len(qo) - 1
    small = n // 8
    d_reads = torch.from_numpy(qb).cuda()
    d_off = torch.from_numpy(qo.view(np.int64)).cuda()
    cap = 16 * n + 64

    def run(flags, nn):
        d_out = torch.zeros(nn * native.READ_RESULT_DTYPE.itemsize, dtype=torch.uint8,
                            device="cuda")
        d_top = torch.zeros(cap * 4, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        ix.classify_device(d_reads, d_off, nn, d_out, d_top, cap, flags=flags)
        out = d_out.cpu().numpy().view(native.READ_RESULT_DTYPE)
        top = d_top.cpu().numpy().view(np.uint32)
        return out["highest"].view(np.uint32).tolist(), native.tie_sets(out, top)

    graph = native.MH_CLASSIFY_BATCHED | native.MH_CLASSIFY_GRAPH
    eager_small = run(native.MH_CLASSIFY_BATCHED, small)
    assert run(graph, small) == eager_small
    eager_all = run(native.MH_CLASSIFY_BATCHED, n)  # grows the arena
    assert run(graph, small) == eager_small
    text = b"".join(b">r%d\n" % i + b"ACGT" * 15 + b"\n" for i in range(400000))
    seq, so, _, _ = h.parse_fasta(torch.from_numpy(np.frombuffer(text, np.uint8).copy()).cuda())
    assert so.numel() == 400001 and seq.numel() == 400000 * 60
    assert run(graph, small) == eager_small
    assert run(graph, n) == eager_all
    ix.free()
    h.close()


def test_graph_not_replayed_for_a_new_index_at_a_reused_address():
    """A captured call holds the index's buffers.  Freeing the index and
    building another (whose handle may come back at the same host address)
    must capture again: the graph call equals the eager call on the new
    index, not the replay of kernels reading the freed one."""
    import torch
    h = native.Handle(0, 4096)
    graph = native.MH_CLASSIFY_BATCHED | native.MH_CLASSIFY_GRAPH
    results = []
    w = D.workload(1, 0.05)
    rb, ro, qb, qo, _ = D.generate(w)
    n = len(qo) - 1
    d_reads = torch.from_numpy(qb).cuda()
    d_off = torch.from_numpy(qo.view(np.int64)).cuda()
    cap = 16 * n + 64
    d_out = torch.zeros(n * native.READ_RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    d_top = torch.zeros(cap * 4, dtype=torch.uint8, device="cuda")
    for variant in range(3):
        # a different reference set each time: the reads match only the first
        rb2 = rb.copy()
        if variant:
            rng = np.random.default_rng(variant)
            rb2[:] = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, len(rb2))]
        ix = h.build_index(rb2, ro, w.threshold)
        per = []
        for flags in (native.MH_CLASSIFY_BATCHED, graph):
            d_out.fill_(0xAB)
            torch.cuda.synchronize()
            ix.classify_device(d_reads, d_off, n, d_out, d_top, cap, flags=flags)
            out = d_out.cpu().numpy().view(native.READ_RESULT_DTYPE)
            top = d_top.cpu().numpy().view(np.uint32)
            per.append((out["highest"].view(np.uint32).tolist(), native.tie_sets(out, top)))
        assert per[0] == per[1], f"variant {variant}: graph call differs from eager"
        results.append(per[0])
        ix.free()
    assert results[0] != results[1]
    h.close()
