"""Parity of the HIP path with the oracle and with the reference's own outputs.

Integer and index work is compared bit for bit; the float Score and relative
score are compared as bit patterns too (the path computes them with the same
IEEE operations and the same libm logf values as the reference).
"""
import numpy as np
import pytest

from mirage_amd import datagen as D
from mirage_amd import native
from oracle import oracle as O
from tests import _data

pytestmark = pytest.mark.gpu


def bits(a):
    return np.asarray(a, dtype=np.float32).view(np.uint32)


def compare_reads(gout, gtop, oout, otop, label=""):
    n = len(oout)
    assert len(gout) == n
    otie = O.top_sets(oout, otop)
    gtie = native.tie_sets(gout, gtop)
    bad = []
    for i in range(n):
        ok = (bits(gout["highest"][i]) == bits(oout["highest"][i]) and
              gout["strand"][i] == oout["strand"][i] and
              gout["n_cand"][i] == oout["n_cand"][i] and gtie[i] == otie[i])
        if not ok:
            bad.append(i)
    if bad:
        i = bad[0]
        pytest.fail(f"{label}: {len(bad)}/{n} reads differ; first read {i}: gpu "
                    f"(h={gout['highest'][i]!r}, s={gout['strand'][i]}, nc={gout['n_cand'][i]}, "
                    f"top={gtie[i]}, st={gout['status'][i]}) oracle (h={oout['highest'][i]!r}, "
                    f"s={oout['strand'][i]}, nc={oout['n_cand'][i]}, top={otie[i]})")


def compare_cands(gc, oc, label=""):
    key = lambda c: (int(c["read"]), int(c["strand"]), int(c["ref"]))  # noqa: E731
    g = sorted(gc.tolist(), key=lambda t: (t[0], t[1], t[2]))
    o = sorted(oc.tolist(), key=lambda t: (t[0], t[1], t[2]))
    assert len(g) == len(o), f"{label}: {len(g)} gpu candidates vs {len(o)} oracle"
    for a, b in zip(g, o):
        # (read, strand, ref, hits, sum_r, sum_c, score, rel, n_seeds, pad)
        fa = (a[:6], np.float32(a[6]).view(np.uint32), np.float32(a[7]).view(np.uint32), a[8])
        fb = (b[:6], np.float32(b[6]).view(np.uint32), np.float32(b[7]).view(np.uint32), b[8])
        assert fa == fb, f"{label}: candidate differs gpu {a} oracle {b}"
    del key


@pytest.mark.parametrize("name", _data.golden_names())
def test_golden_fixture(handle, name):
    """Reference binary output == GPU output for every read of the fixture."""
    fx, rb, ro, qb, qo = _data.load_golden(name)
    ix = handle.build_index(rb, ro, fx["threshold"])
    out, top, cands = ix.classify_host(qb, qo, want_cands=True)
    exp = _data.expected_lines(fx, len(qo) - 1)
    ties = native.tie_sets(out, top)
    for i, e in enumerate(exp):
        got = None if out["n_top"][i] == 0 else ("%f" % out["highest"][i], ties[i])
        assert got == e, f"{name} read {i}: gpu {got} reference {e}"
    oix = O.Index(rb, ro, fx["threshold"])
    oout, otop, ocands = oix.classify(qb, qo, want_cands=True)
    compare_reads(out, top, oout, otop, name)
    compare_cands(cands, ocands, name)


WORKLOADS = [(1, 0.2), (2, 0.004), (3, 0.0005), (5, 0.0005)]


@pytest.fixture(scope="module")
def workloads():
    cache = {}

    def get(num, scale):
        key = (num, scale)
        if key not in cache:
            w = D.workload(num, scale)
            if num == 5:
                w.refs.total = 12_000_000
            if num == 3:
                w.refs.n = 300  # keep the oracle index small
            cache[key] = (w, D.generate(w))
        return cache[key]
    return get


@pytest.mark.parametrize("num,scale", WORKLOADS)
@pytest.mark.parametrize("flags", [0, native.MH_CLASSIFY_BATCHED,
                                   native.MH_CLASSIFY_BATCHED | native.MH_CLASSIFY_PREPROBE])
def test_workload_matches_oracle(handle, workloads, num, scale, flags):
    w, (rb, ro, qb, qo, truth) = workloads(num, scale)
    ix = handle.build_index(rb, ro, w.threshold)
    out, top, cands = ix.classify_host(qb, qo, flags=flags, want_cands=True)
    oix = O.Index(rb, ro, w.threshold)
    oout, otop, ocands = oix.classify(qb, qo, want_cands=True)
    compare_reads(out, top, oout, otop, f"workload {num}")
    compare_cands(cands, ocands, f"workload {num}")


def test_direct_hit_path_is_exact(handle, workloads):
    """The direct-hit shortcut gives the same results as the full path, and
    it is taken for the exact-substring reads of workload 2."""
    w, (rb, ro, qb, qo, truth) = workloads(2, 0.004)
    ix = handle.build_index(rb, ro, w.threshold)
    a, ta, ca = ix.classify_host(qb, qo, want_cands=True)
    direct = handle.counters()["direct"]
    b, tb, cb = ix.classify_host(qb, qo, flags=native.MH_CLASSIFY_NO_DIRECT, want_cands=True)
    assert handle.counters()["direct"] == 0
    assert direct > 0
    assert np.array_equal(bits(a["highest"]), bits(b["highest"]))
    assert native.tie_sets(a, ta) == native.tie_sets(b, tb)
    compare_cands(ca, cb, "direct vs full")


def test_index_buckets_match_oracle(handle):
    w = D.workload(1)
    rb, ro = D.make_refs(w.refs, w.seed)
    ix = handle.build_index(rb, ro, w.threshold)
    oix = O.Index(rb, ro, w.threshold)
    info = ix.info()
    assert info["n_entries"] == oix.num_entries
    assert info["n_buckets"] + (1 if info["empty_hash"] != 0xFFFFFFFF else 0) == oix.num_buckets
    rng = np.random.default_rng(7)
    k = w.threshold - w.threshold // 2 + 1
    # hashes of real windows (present) and random hashes (mostly absent)
    starts = rng.integers(0, len(rb) - k, 300)
    hs = [O.hash28(rb[s:s + k].tobytes()) for s in starts] + rng.integers(0, 1 << 28, 300).tolist()
    hs.append(info["empty_hash"])
    for h in hs:
        if h == 0xFFFFFFFF:
            continue
        assert ix.bucket(int(h)) == oix.bucket(int(h)), f"bucket {h:#x}"


@pytest.mark.parametrize("name", ["t30_gc_lognormal", "t31_mixed_indels"])
@pytest.mark.parametrize("flags", [0, native.MH_CLASSIFY_BATCHED])
def test_ultrafast_matches_oracle(handle, name, flags):
    """UltraFast mode (MatrixCPU.cpp:69: no head/tail extension, a seed
    counts its own length only) against the oracle's, per read and per
    candidate."""
    fx, rb, ro, qb, qo = _data.load_golden(name)
    ix = handle.build_index(rb, ro, fx["threshold"])
    out, top, cands = ix.classify_host(qb, qo, flags=flags | native.MH_CLASSIFY_ULTRAFAST,
                                       want_cands=True)
    oout, otop, ocands = O.Index(rb, ro, fx["threshold"]).classify(qb, qo, want_cands=True,
                                                                   ultrafast=True)
    compare_reads(out, top, oout, otop, f"{name} ultrafast")
    compare_cands(cands, ocands, f"{name} ultrafast")
    # the mode matters on these reads: extension changes some scores
    eout, _, _ = O.Index(rb, ro, fx["threshold"]).classify(qb, qo)
    assert not np.array_equal(bits(eout["highest"]), bits(oout["highest"]))
    assert handle.counters()["ext_compares"] == 0
    ix.free()
