"""MirAge's result logs, restated in Python (test infrastructure).

For each classified read the reference prints one line per group of equal
relative scores (only the highest group with output_firstresults_only: 1,
the default) into four files (MirAge/Analyser.cpp:378-491, line format
:226-376, files LogManager.cpp:294-404):

  results_Raw.log              every line
  results_HighSensitivity.log  lines whose score >= 0.4 (LogManager.cpp:10)
  results_HighPrecision.log    lines whose score >= 1 (:11)
  results_Balanced.log         lines whose score >= 0.7 (:12)

each starting with the header line
"InputID;InputSequence;Confident;TaxIDs;Occurence_Per_TaxID;Headers".
A line is "InputID;read header;score;taxonomy ids;occurrences;reference
headers", headers without their '>'-prefixed name's trailing NUL, the score
as C++ to_string(float) ("%f").

The references of a group come out in the order two library sorts leave
them: std::sort of the read's scored candidates by score, descending
(:404), then std::stable_sort by taxonomy occurrence with a comparator that
puts taxonomy id -1 first whatever the other element is (:266-293).
Neither is a plain ordering (std::sort is not stable; the comparator is not
a strict weak ordering), so both are called in the library itself, built
into the C++ oracle with the reference's comparators (mo_std_sort_desc,
mo_std_stable_sort_tax in oracle/mirage_oracle.cpp; GNU libstdc++ of the
image's GCC 11, the reference's toolchain).  A read's scored candidates are
those with Score > -1 on its surviving strand, in candidate order, i.e.
ascending reference id (MatrixCPU.cpp:323-370).
"""
from collections import Counter

import numpy as np

from . import oracle as O

KINDS = ("Raw", "HighSensitivity", "HighPrecision", "Balanced")
HEADER = "InputID;InputSequence;Confident;TaxIDs;Occurence_Per_TaxID;Headers\n"
THRESHOLDS = {"HighSensitivity": np.float32(0.4), "HighPrecision": np.float32(1.0),
              "Balanced": np.float32(0.7)}


def to_string(x):
    """std::to_string(float): vsnprintf "%f" of the value as a double."""
    return "%f" % float(np.float32(x))


def sort_by_score_desc(scores):
    """std::sort of [(rel, ref)] by rel, descending (Analyser.cpp:404)."""
    return O.std_sort_desc(scores)


def stable_sort_by_taxonomy(refs, taxids):
    """std::stable_sort of reference ids with the comparator of
    Analyser.cpp:268-293 (taxonomy id -1 first; then higher occurrence)."""
    return O.std_stable_sort_tax(refs, taxids)


def score_groups(scores, first_only=True):
    """scores: [(rel as float32, ref)] in candidate order.  Returns
    [(rel, [refs])] as LogMatrixResult groups them (Analyser.cpp:401-424)."""
    a = [(np.float32(r), int(i)) for r, i in sort_by_score_desc(list(scores))]
    out = [(a[0][0], [])]
    for rel, ref in a:
        if rel != out[-1][0]:
            if first_only:
                break
            out.append((rel, []))
        out[-1][1].append(ref)
    return out


def line(input_id, read_header, rel, refs, ref_headers, taxids):
    """LogResultsToFileManager (Analyser.cpp:226-376) for one group."""
    s = "%d;%s;%s;" % (input_id, read_header, to_string(rel))
    if not refs:
        return s + "?;"
    tax = [taxids[r] for r in refs]
    occ = Counter(tax)
    res = [int(r) for r in stable_sort_by_taxonomy(list(refs), taxids)]
    cur = taxids[res[0]]
    max_occ = occ[cur]
    for column in ("tax", "occ"):  # :299-351; cur carries over between them
        for i, r in enumerate(res):
            t = taxids[r]
            if i > 0 and t == cur:
                continue
            if occ[t] != max_occ:
                break
            cur = t
            if i != 0:
                s += ","
            s += str(cur) if column == "tax" else str(occ[t])
        s += ";"
    for i, r in enumerate(res):  # :354-370
        if occ[taxids[r]] != max_occ:
            break
        if i != 0:
            s += ","
        s += ref_headers[r]
    return s


def logs(reads, ref_headers, taxids=None, first_only=True):
    """The four files as text.  reads: [(input_id, read header, highest,
    scores)] in output order, scores as for score_groups; ref_headers
    without their trailing NUL."""
    if taxids is None:
        taxids = [-1] * len(ref_headers)
    out = {k: HEADER for k in KINDS}
    for input_id, header, highest, scores in reads:
        if not scores or np.float32(highest) == 0:  # Analyser.cpp:388
            continue
        for rel, refs in score_groups(scores, first_only):
            text = line(input_id, header, rel, refs, ref_headers, taxids) + "\n"
            for k in KINDS:
                if k == "Raw" or np.float32(rel) >= THRESHOLDS[k]:
                    out[k] += text
    return out


def reads_from_oracle(out, cands, read_headers, first_read_id=0):
    """Per read (input id, header, highest, scores) from a classification:
    the scored candidates (score > -1) of the surviving strand in ascending
    reference order."""
    by_read = {}
    for c in cands:
        if c["score"] > -1:
            by_read.setdefault((int(c["read"]), int(c["strand"])), []).append(
                (np.float32(c["rel"]), int(c["ref"])))
    rows = []
    for i in range(len(out)):
        if out["n_top"][i] == 0:
            continue
        sc = sorted(by_read.get((i, int(out["strand"][i])), []), key=lambda t: t[1])
        rows.append((first_read_id + i, read_headers[i], np.float32(out["highest"][i]), sc))
    return rows
