"""Write end-to-end golden fixtures with the reference's own CPU binary.

Run in the build container only (the reference is not on the GPU box):

    make -C oracle ref            # builds oracle/_ref/mirage_ref from /root/reference
    python tests/golden/make_ref_goldens.py

Each fixture is seeded synthetic data (references and reads as strings) plus
what the reference printed for it: per read id, the relative score string and
the top-tie reference ids (results_Raw.log, Analyser.cpp:228-376).  Reads
without a line are unclassified.  The binary is run twice per dataset; a
fixture is written only when both runs agree.
"""
import gzip
import json
import os
import shutil
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from mirage_amd import datagen as D  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "mirage_ref")
RUNS = os.path.join(ROOT, "oracle", "_ref", "runs")
CFG_TEMPLATE = "/root/reference/MirAge/data/mirage.identifier.cfg"
OUT_DIR = os.path.dirname(os.path.abspath(__file__))
ASAN = ("detect_leaks=0:allow_user_segv_handler=0:halt_on_error=0:alloc_dealloc_mismatch=0:"
        "new_delete_type_mismatch=0:detect_container_overflow=0")


def datasets():
    """(name, T, refs spec, reads spec, seed, edits) for each fixture."""
    L = D.LengthSpec
    yield ("t22_uniform", 22,
           D.RefSpec(n=24, lengths=L(mode=0, a=6000)),
           D.ReadSpec(n=400, lengths=L(mode=0, a=150), p_random=0.1, sub_rate=0.01, rc_frac=0.5),
           101, None)
    yield ("t30_gc_lognormal", 30,
           D.RefSpec(n=40, lengths=L(mode=1, a=8000, b=0.8, lo=600, cap=60000), gc_lo=0.4,
                     gc_hi=0.6),
           D.ReadSpec(n=500, lengths=L(mode=0, a=150), p_exact=0.05, p_random=0.1,
                      sub_rate=0.02, rc_frac=0.5),
           102, None)
    yield ("t30_mixed_indels", 30,
           D.RefSpec(n=30, lengths=L(mode=1, a=10000, b=0.7, lo=2500, cap=60000), gc_lo=0.4,
                     gc_hi=0.6),
           D.ReadSpec(n=300, lengths=L(mode=3, a=75, b=300, c=900, d=0.6, mix=0.7, lo=75,
                                       cap=2400),
                      sub_rate=0.03, indel_rate=0.005, rc_frac=0.5),
           103, None)
    yield ("t38_repetitive", 38,
           D.RefSpec(n=40, lengths=L(mode=1, a=9000, b=0.5, lo=3000, cap=40000), seg_pool=6,
                     seg_lo=300, seg_hi=1500, seg_frac=0.6),
           D.ReadSpec(n=400, lengths=L(mode=0, a=150), sub_rate=0.01, rc_frac=0.5),
           104, None)
    yield ("t18_fragment_rule", 18,
           D.RefSpec(n=20, lengths=L(mode=0, a=5000), gc_lo=0.3, gc_hi=0.7),
           D.ReadSpec(n=300, lengths=L(mode=0, a=120), p_random=0.05, sub_rate=0.01,
                      rc_frac=0.5),
           105, None)
    yield ("t30_n_runs", 30,
           D.RefSpec(n=16, lengths=L(mode=0, a=8000)),
           D.ReadSpec(n=300, lengths=L(mode=0, a=150), sub_rate=0.01, rc_frac=0.5),
           106, "n_runs")
    yield ("t22_tiny_mixed", 22,
           D.RefSpec(n=8, lengths=L(mode=2, a=300, b=3000)),
           D.ReadSpec(n=250, lengths=L(mode=2, a=40, b=400), p_exact=0.2, p_random=0.1,
                      sub_rate=0.01, rc_frac=0.5),
           107, None)


def apply_edits(kind, rb, ro, qb, qo, rng):
    if kind != "n_runs":
        return
    # N runs inside some references (window cutoff, Hashmap.cpp:2882-2895) and
    # a few N bases inside reads.
    for r in range(0, len(ro) - 1, 3):
        a, b = int(ro[r]), int(ro[r + 1])
        at = a + int(rng.integers(200, b - a - 200))
        rb[at:at + int(rng.integers(5, 40))] = ord("N")
    for i in range(0, len(qo) - 1, 11):
        a = int(qo[i])
        rb_at = a + int(rng.integers(0, int(qo[i + 1]) - a))
        qb[rb_at] = ord("N")


def write_fasta(path, prefix, by, off):
    with open(path, "wb") as f:
        for i in range(len(off) - 1):
            f.write(b">%s%d\n" % (prefix.encode(), i))
            s = by[int(off[i]):int(off[i + 1])].tobytes()
            for j in range(0, len(s), 80):
                f.write(s[j:j + 80] + b"\n")


def run_reference(workdir, threshold):
    if os.path.exists(workdir):
        shutil.rmtree(workdir)
    os.makedirs(os.path.join(workdir, "mirage", "data"))
    os.makedirs(os.path.join(workdir, "work"))
    os.makedirs(os.path.join(workdir, "out"))
    cfg = open(CFG_TEMPLATE).read().splitlines()
    lines = []
    for ln in cfg:
        key = ln.split(":", 1)[0].strip().lower()
        if key == "query":
            ln = "Query:" + os.path.join(workdir, "..", "query.fasta")
        elif key == "reference":
            ln = "Reference:" + os.path.join(workdir, "..", "reference.fasta")
        elif key == "outputpath":
            ln = "OutputPath:" + os.path.join(workdir, "out")
        elif key == "considerthreshold":
            ln = "considerthreshold:%d" % threshold
        lines.append(ln)
    with open(os.path.join(workdir, "mirage", "data", "mirage.identifier.cfg"), "w") as f:
        f.write("\n".join(lines) + "\n")
    env = dict(os.environ, ASAN_OPTIONS=ASAN)
    with open(os.path.join(workdir, "run.log"), "wb") as log:
        subprocess.run(["timeout", "900", REF_BIN], cwd=os.path.join(workdir, "work"), env=env,
                       stdout=log, stderr=subprocess.STDOUT, check=True)
    outs = []
    for root, _, files in os.walk(os.path.join(workdir, "out")):
        for fn in files:
            if fn == "results_Raw.log":
                outs.append(os.path.join(root, fn))
    assert len(outs) == 1, outs
    res = {}
    with open(outs[0]) as f:
        next(f)
        for line in f:
            line = line.rstrip("\n")
            if not line:
                continue
            cols = line.split(";")
            rid = int(cols[0])
            refs = sorted(int(h.lstrip(">")[1:]) for h in cols[5].split(","))
            res[rid] = [cols[2], refs]
    return res


def main():
    only = set(sys.argv[1:])
    for name, T, rspec, qspec, seed, edits in datasets():
        if only and name not in only:
            continue
        rb, ro = D.make_refs(rspec, seed)
        qb, qo, truth = D.make_reads(qspec, rb, ro, seed)
        apply_edits(edits, rb, ro, qb, qo, np.random.default_rng(seed))
        base = os.path.join(RUNS, name)
        os.makedirs(base, exist_ok=True)
        write_fasta(os.path.join(base, "reference.fasta"), "r", rb, ro)
        write_fasta(os.path.join(base, "query.fasta"), "q", qb, qo)
        r1 = run_reference(os.path.join(base, "a"), T)
        r2 = run_reference(os.path.join(base, "b"), T)
        if r1 != r2:
            print(f"{name}: reference runs disagree, fixture not written")
            continue
        fx = {
            "name": name,
            "threshold": T,
            "source": "oracle/_ref/mirage_ref (reference CPU build, ASan -O0), two agreeing runs",
            "refs": [rb[int(ro[i]):int(ro[i + 1])].tobytes().decode() for i in range(len(ro) - 1)],
            "reads": [qb[int(qo[i]):int(qo[i + 1])].tobytes().decode()
                      for i in range(len(qo) - 1)],
            "expected": {str(k): v for k, v in sorted(r1.items())},
        }
        with gzip.open(os.path.join(OUT_DIR, f"ref_{name}.json.gz"), "wt") as f:
            json.dump(fx, f)
        print(f"{name}: {len(r1)} classified of {len(qo) - 1}")


if __name__ == "__main__":
    main()
