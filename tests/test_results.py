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
    orig_sort, orig_stable = R.sort_by_score_desc, R.stable_sort_by_taxonomy
    try:
        R.sort_by_score_desc = lambda a: sorted(a, key=lambda t: -t[0])
        assert oracle_logs(fx, out, cands)["Raw"] != fx["logs"]["Raw"]
        R.sort_by_score_desc = orig_sort
        R.stable_sort_by_taxonomy = lambda refs, taxids: list(refs)
        assert oracle_logs(fx, out, cands)["Raw"] != fx["logs"]["Raw"]
    finally:
        R.sort_by_score_desc, R.stable_sort_by_taxonomy = orig_sort, orig_stable


def test_shipped_sample_line():
    """The reference's own sample data (MirAge/data/reference.fasta and
    query.fasta) at its shipped T = 31 (mirage.identifier.cfg:22): the oracle
    prints the line the reference binary printed, read 1 unclassified."""
    fx, n, n_refs, out, top, cands = classified("t31_shipped_data")
    assert fx["threshold"] == 31 and n == 2 and n_refs == 2
    logs = oracle_logs(fx, out, cands)
    assert logs["Raw"] == R.HEADER + "0;>RM2|S2|R14/1;0.627418;-1;2;>2,>1\n"
    assert out["n_top"][1] == 0


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
