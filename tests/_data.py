"""Shared test data: golden fixtures and seeded workloads."""
import glob
import gzip
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def pack(seqs):
    off = np.zeros(len(seqs) + 1, np.uint64)
    off[1:] = np.cumsum([len(s) for s in seqs])
    by = np.frombuffer("".join(seqs).encode(), np.uint8).copy() if seqs else np.zeros(0, np.uint8)
    return by, off


def golden_names():
    return sorted(os.path.basename(p)[4:-8] for p in glob.glob(os.path.join(GOLDEN, "ref_*.json.gz")))


def load_golden(name):
    with gzip.open(os.path.join(GOLDEN, f"ref_{name}.json.gz"), "rt") as f:
        fx = json.load(f)
    if "refs" not in fx:  # only the FASTA texts are stored: parse them
        from oracle import fasta as F
        fx["refs"] = [x.decode("latin-1") for x in
                      F.parse(fx["reference_fasta"].encode("latin-1"))[0]]
        fx["reads"] = [x.decode("latin-1") for x in
                       F.parse(fx["query_fasta"].encode("latin-1"))[0]]
    rb, ro = pack(fx["refs"])
    qb, qo = pack(fx["reads"])
    return fx, rb, ro, qb, qo


def expected_lines(fx, n_reads):
    """Per read: None (no line printed) or (rel string, sorted ref ids)."""
    out = []
    for i in range(n_reads):
        e = fx["expected"].get(str(i))
        out.append(None if e is None else (e[0], tuple(e[1])))
    return out


def mhm_names():
    """Datasets whose reference-written .mhm file is a fixture."""
    return sorted(os.path.basename(p)[4:-7] for p in glob.glob(os.path.join(GOLDEN, "ref_*.mhm.gz")))


def mhm_bytes(name):
    with gzip.open(os.path.join(GOLDEN, f"ref_{name}.mhm.gz"), "rb") as f:
        return f.read()


def mhm_file(name, directory):
    """Writes the decompressed .mhm fixture under directory; returns its path."""
    path = os.path.join(str(directory), f"{name}.mhm")
    with open(path, "wb") as f:
        f.write(mhm_bytes(name))
    return path


def headers(n):
    """Reference headers as MirAge keeps them for the fixtures' FASTA names."""
    return [b">r%d\0" % i for i in range(n)]


def fasta_headers(fx):
    """(read headers, reference headers) as MirAge keeps them ('>' name
    NUL): parsed from the dataset's FASTA texts when the fixture holds them,
    else the generator's '>q<i>' and '>r<i>'."""
    if "reference_fasta" in fx:
        from oracle import fasta as F
        return (F.parse(fx["query_fasta"].encode("latin-1"))[1],
                F.parse(fx["reference_fasta"].encode("latin-1"))[1])
    return ([b">q%d\0" % i for i in range(len(fx["reads"]))],
            [b">r%d\0" % i for i in range(len(fx["refs"]))])
