"""Serving loop and autotuner on the GPU: results are the same alone and in
the mix, and tile choices do not change results."""
import numpy as np
import pytest

from mirage_amd import datagen as D
from mirage_amd import native

pytestmark = pytest.mark.gpu


def split(qb, qo, sizes):
    reqs, r = [], 0
    for n in sizes:
        a, b = int(qo[r]), int(qo[r + n])
        reqs.append((qb[a:b], qo[r:r + n + 1] - qo[r]))
        r += n
    return reqs


@pytest.fixture(scope="module")
def served(handle):
    w = D.workload(4, 0.005)  # 100 refs ~10 Mbp, reads 100-5000 bp
    rb, ro, qb, qo, _ = D.generate(w)
    ix = handle.build_index(rb, ro, w.threshold)
    return w, ix, qb, qo


def test_each_request_same_alone_and_in_mix(served):
    w, ix, qb, qo = served
    rng = np.random.default_rng(1)
    sizes = []
    left = len(qo) - 1
    while left > 0:
        n = int(min(left, max(1, 10 ** rng.uniform(0, 2.7))))
        sizes.append(n)
        left -= n
    reqs = split(qb, qo, sizes)
    alone = [ix.classify_host(b, o, flags=native.MH_CLASSIFY_BATCHED) for b, o in reqs]
    srv = native.Server(ix, max_inflight=3, max_batch_reads=1024)
    ids = [srv.submit(b, o) for b, o in reqs]
    for (aout, atop, _), rid in zip(alone, ids):
        out, top, info = srv.wait(rid)
        assert np.array_equal(out["highest"].view(np.uint32), aout["highest"].view(np.uint32))
        assert np.array_equal(out["strand"], aout["strand"])
        assert np.array_equal(out["n_cand"], aout["n_cand"])
        assert native.tie_sets(out, top) == native.tie_sets(aout, atop)
        assert info["submit_s"] <= info["admit_s"] <= info["done_s"]
    srv.stop()


def test_server_admits_while_busy(served):
    """Requests submitted while batches run are admitted into later batches."""
    w, ix, qb, qo = served
    reqs = split(qb, qo, [25] * 20)
    srv = native.Server(ix, max_inflight=2, max_batch_reads=50)
    ids = [srv.submit(b, o) for b, o in reqs]
    infos = [srv.wait(r)[2] for r in ids]
    srv.stop()
    admits = sorted(i["admit_s"] for i in infos)
    assert admits[-1] > admits[0]  # not all in one batch


def test_served_reads_at_tile_boundaries(served, handle):
    """The server launches only the length tiles its batch's longest read
    can reach: single-read batches of every tile's maximum length and one
    base more give the results of a batched call, which launches every
    tile."""
    w, ix, qb, qo = served
    tiles = handle.get_tiles(ix)[0]
    lens = np.diff(qo)
    src = int(np.argmax(lens))
    a, longest = int(qo[src]), int(lens[src])
    cut = sorted({m + d for t in tiles[:-1] for m in [t[0]] for d in (0, 1)
                  if 0 < m + d <= longest} | {1, longest})
    assert len(cut) >= 4
    reqs = [(qb[a:a + n].copy(), np.array([0, n], dtype=qo.dtype)) for n in cut]
    srv = native.Server(ix, max_inflight=1, max_batch_reads=1)
    for b, o in reqs:
        aout, atop, _ = ix.classify_host(b, o, flags=native.MH_CLASSIFY_BATCHED)
        out, top, info = srv.wait(srv.submit(b, o))
        assert info["n_reads"] == 1
        assert np.array_equal(out["highest"].view(np.uint32), aout["highest"].view(np.uint32)), len(b)
        assert np.array_equal(out["strand"], aout["strand"]), len(b)
        assert native.tie_sets(out, top) == native.tie_sets(aout, atop), len(b)
    srv.stop()


def test_autotune_keeps_results(handle):
    import torch
    w = D.workload(3, 0.001)
    w.refs.n = 200
    rb, ro, qb, qo, _ = D.generate(w)
    ix = handle.build_index(rb, ro, w.threshold)
    before, btop, _ = ix.classify_host(qb, qo, flags=native.MH_CLASSIFY_BATCHED)
    defaults = handle.get_tiles(ix)[0]
    d_reads = torch.from_numpy(qb).cuda()
    d_off = torch.from_numpy(qo.view(np.int64)).cuda()
    torch.cuda.synchronize()
    tiles = handle.autotune(ix, d_reads, d_off, len(qo) - 1, iters=1)
    assert len(tiles) == len(defaults)
    after, atop, _ = ix.classify_host(qb, qo, flags=native.MH_CLASSIFY_BATCHED)
    handle.set_tiles(None)
    assert np.array_equal(before["highest"].view(np.uint32), after["highest"].view(np.uint32))
    assert native.tie_sets(before, btop) == native.tie_sets(after, atop)
