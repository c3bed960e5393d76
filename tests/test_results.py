"""The reference's result logs: the oracle's restatement and the library's
writer, both checked against the four log files the reference binary wrote
for every golden dataset (tests/golden, key "logs").

No GPU: the writer is host code, fed here with the oracle's classification
(which the GPU path matches bit for bit elsewhere).
"""
import numpy as np
import pytest

from mirage_amd import native
from oracle import oracle as O
from oracle import results as R
from oracle import stdsort
from tests import _data

NAMES = _data.golden_names()


def headers(prefix, n):
    return [b">%s%d\0" % (prefix.encode(), i) for i in range(n)]


def classified(name):
    fx, rb, ro, qb, qo = _data.load_golden(name)
    out, top, cands = O.Index(rb, ro, fx["threshold"]).classify(qb, qo, want_cands=True)
    return fx, len(qo) - 1, len(ro) - 1, out, top, cands


def printed(hs):
    return [h[:-1].decode("latin-1") for h in hs]


def oracle_logs(fx, out, cands, first_only=True, taxids=None):
    qh, rh = _data.fasta_headers(fx)
    rows = R.reads_from_oracle(out, cands, printed(qh))
    return R.logs(rows, printed(rh), taxids, first_only)


def top_pool(out, top):
    """The oracle's results as the library lays them out: mh_read_result
    records and a top-tie pool with offsets."""
    pool = []
    res = np.zeros(len(out), native.READ_RESULT_DTYPE)
    for f in ("highest", "strand", "n_top", "n_cand"):
        res[f] = out[f]
    for i in range(len(out)):
        res["top_off"][i] = len(pool)
        pool.extend(top[i, :out["n_top"][i]].tolist())
    return res, np.array(pool or [0], np.uint32)


def written(tmp_path):
    return {k: (tmp_path / f"results_{k}.log").read_text() for k in R.KINDS}


def test_stdsort_matches_libstdcxx():
    rng = np.random.default_rng(1)
    for _ in range(1500):
        n, k = int(rng.integers(0, 120)), int(rng.integers(1, 6))
        sc = [(np.float32(rng.integers(0, k)) / 4, i) for i in range(n)]
        a = list(sc)
        stdsort.sort(a, lambda l, r: l[0] > r[0])
        assert [(float(x), y) for x, y in a] == [(float(x), y) for x, y in O.std_sort_desc(sc)]
        nt = int(rng.integers(1, 40))
        tax = rng.integers(-1, 4, nt).tolist()
        ids = rng.permutation(nt)[:int(rng.integers(0, nt + 1))].tolist()
        occ = {t: sum(tax[r] == t for r in ids) for t in set(tax)}

        def comp(a_, b_):
            ta, tb = tax[a_], tax[b_]
            if ta == -1:
                return True
            if tb == -1:
                return False
            if ta == tb or occ[ta] == occ[tb]:
                return False
            return occ[ta] > occ[tb]
        c = list(ids)
        stdsort.stable_sort(c, comp)
        assert c == O.std_stable_sort_tax(ids, tax)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_logs_match_reference(name):
    fx, n, n_refs, out, top, cands = classified(name)
    logs = oracle_logs(fx, out, cands)
    for k in R.KINDS:
        assert logs[k] == fx["logs"][k], (name, k)


def test_planted_sort_mistakes_are_caught():
    """A stable score sort, or a stable taxonomy sort with a valid
    comparator, changes the reference order of tied references."""
    fx, n, n_refs, out, top, cands = classified("t30_ties20")
    orig_sort, orig_stable = stdsort.sort, stdsort.stable_sort
    try:
        stdsort.sort = lambda a, comp: a.sort(key=lambda t: -t[0])
        assert oracle_logs(fx, out, cands)["Raw"] != fx["logs"]["Raw"]
        stdsort.sort = orig_sort
        stdsort.stable_sort = lambda a, comp: None
        assert oracle_logs(fx, out, cands)["Raw"] != fx["logs"]["Raw"]
    finally:
        stdsort.sort, stdsort.stable_sort = orig_sort, orig_stable


@pytest.mark.parametrize("name", NAMES)
def test_writer_matches_reference(name, tmp_path):
    """mh_write_results with the candidate output, and without it (the top
    ties stand in, exact here: no read has more than 16 scored candidates
    unless they all tie), gives the reference's files byte for byte."""
    fx, n, n_refs, out, top, cands = classified(name)
    res, pool = top_pool(out, top)
    qh, rh = _data.fasta_headers(fx)
    for with_cands in (True, False):
        native.write_results(tmp_path, res, pool, qh, rh, cands=cands if with_cands else None)
        got = written(tmp_path)
        for k in R.KINDS:
            assert got[k] == fx["logs"][k], (name, k, with_cands)


def test_writer_append_groups_taxids_thresholds(tmp_path):
    """Beyond the reference's own files: a second batch appended (InputIDs
    continue), a line per score group, taxonomy ids and other thresholds,
    against the restatement."""
    fx, n, n_refs, out, top, cands = classified("t38_repetitive")
    res, pool = top_pool(out, top)
    half = n // 2
    sel = cands["read"] < half
    c2 = cands[~sel].copy()
    c2["read"] -= half
    _, pool1 = top_pool(out[:half], top[:half])
    res2, pool2 = top_pool(out[half:], top[half:])
    native.write_results(tmp_path, top_pool(out[:half], top[:half])[0], pool1,
                         headers("q", n)[:half], headers("r", n_refs), cands=cands[sel])
    native.write_results(tmp_path, res2, pool2, headers("q", n)[half:], headers("r", n_refs),
                         cands=c2, first_read_id=half, append=True)
    got = written(tmp_path)
    for k in R.KINDS:
        assert got[k] == fx["logs"][k], k
    rng = np.random.default_rng(5)
    taxids = rng.integers(-1, 5, n_refs).tolist()
    native.write_results(tmp_path, res, pool, headers("q", n), headers("r", n_refs), cands=cands,
                         taxids=taxids, all_groups=True, thresholds=(0.5, 0.95, 0.8))
    got = written(tmp_path)
    rows = R.reads_from_oracle(out, cands, [">q%d" % i for i in range(n)])
    saved = dict(R.THRESHOLDS)
    try:
        R.THRESHOLDS.update({"HighSensitivity": np.float32(0.5),
                             "HighPrecision": np.float32(0.95), "Balanced": np.float32(0.8)})
        exp = R.logs(rows, [">r%d" % i for i in range(n_refs)], taxids, first_only=False)
    finally:
        R.THRESHOLDS.update(saved)
    for k in R.KINDS:
        assert got[k] == exp[k], k
    assert got["Raw"].count("\n") > fx["logs"]["Raw"].count("\n")  # several groups per read
