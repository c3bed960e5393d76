"""Write end-to-end golden fixtures with the reference's own CPU binary.

Run in the build container only (the reference is not on the GPU box):

    make -C oracle ref            # builds oracle/_ref/mirage_ref from /root/reference
    python tests/golden/make_ref_goldens.py

Each fixture is seeded synthetic data (references and reads as strings) plus
what the reference printed for it: per read id, the relative score string and
the top-tie reference ids (results_Raw.log, Analyser.cpp:228-376).  Reads
without a line are unclassified.  The binary is run twice per dataset; a
fixture is written only when both runs agree (results and the .mhm file).
For the datasets in MHM_FIXTURES the reference's .mhm hashmap file is kept
as well (ref_<name>.mhm.gz).
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
    yield ("t30_rc_copies", 30,
           D.RefSpec(n=24, lengths=L(mode=0, a=6000)),
           D.ReadSpec(n=400, lengths=L(mode=0, a=150), sub_rate=0.01, rc_frac=0.5),
           108, "rc_copies")
    yield ("t22_low_complexity", 22,
           D.RefSpec(n=20, lengths=L(mode=0, a=6000)),
           D.ReadSpec(n=400, lengths=L(mode=0, a=150), sub_rate=0.005, rc_frac=0.5),
           109, "low_complexity")
    yield ("t30_variant_copies", 30,
           D.RefSpec(n=30, lengths=L(mode=0, a=5000)),
           D.ReadSpec(n=400, lengths=L(mode=2, a=100, b=400), sub_rate=0.01, rc_frac=0.5),
           110, "variant_copies")
    yield ("t30_engineered", 30, None, None, 111, "engineered")
    yield ("t30_ties20", 30, None, None, 115, "ties20")
    yield ("t22_fasta_forms", 22,
           D.RefSpec(n=6, lengths=L(mode=2, a=300, b=2500)),
           D.ReadSpec(n=150, lengths=L(mode=2, a=40, b=400), p_exact=0.2, p_random=0.1,
                      sub_rate=0.01, rc_frac=0.5),
           116, "fasta_forms")
    yield ("t22_tiny_mixed", 22,
           D.RefSpec(n=8, lengths=L(mode=2, a=300, b=3000)),
           D.ReadSpec(n=250, lengths=L(mode=2, a=40, b=400), p_exact=0.2, p_random=0.1,
                      sub_rate=0.01, rc_frac=0.5),
           107, None)
    yield ("t18_tiny_mixed", 18,
           D.RefSpec(n=6, lengths=L(mode=2, a=300, b=2500), gc_lo=0.3, gc_hi=0.7),
           D.ReadSpec(n=200, lengths=L(mode=2, a=40, b=300), p_exact=0.2, p_random=0.1,
                      sub_rate=0.01, rc_frac=0.5),
           112, None)
    # Odd T (k = s + 2): the reference's shipped default is T = 31
    # (MirAge/data/mirage.identifier.cfg:22).
    yield ("t31_uniform", 31,
           D.RefSpec(n=24, lengths=L(mode=0, a=6000)),
           D.ReadSpec(n=400, lengths=L(mode=0, a=150), p_exact=0.05, p_random=0.1,
                      sub_rate=0.01, rc_frac=0.5),
           117, None)
    yield ("t31_mixed_indels", 31,
           D.RefSpec(n=30, lengths=L(mode=1, a=8000, b=0.7, lo=2000, cap=50000), gc_lo=0.4,
                     gc_hi=0.6, seg_pool=4, seg_lo=300, seg_hi=1200, seg_frac=0.4),
           D.ReadSpec(n=300, lengths=L(mode=3, a=60, b=300, c=900, d=0.6, mix=0.7, lo=60,
                                       cap=2400),
                      sub_rate=0.02, indel_rate=0.005, rc_frac=0.5),
           118, None)
    # The reference's own sample data (MirAge/data/reference.fasta and
    # query.fasta) at its shipped T = 31.
    yield ("t31_shipped_data", 31, None, None, 0, "shipped")
    yield ("t30_tiny_nruns", 30,
           D.RefSpec(n=6, lengths=L(mode=0, a=2000)),
           D.ReadSpec(n=200, lengths=L(mode=0, a=150), sub_rate=0.01, rc_frac=0.5),
           114, "n_runs")




# Datasets whose .mhm file (the reference's own hashmap database, written by
# SaveHashmapTableV3 and loaded back before classification) is kept as a
# fixture too: ref_<name>.mhm.gz.
MHM_FIXTURES = ("t22_tiny_mixed", "t18_tiny_mixed", "t30_tiny_nruns", "t22_fasta_forms")
LOG_KINDS = ("Raw", "HighSensitivity", "HighPrecision", "Balanced")


COMP = np.zeros(256, np.uint8)
for a, b in zip(b"ACGTN", b"TGCAN"):
    COMP[a] = b


def apply_edits(kind, rb, ro, qb, qo, rng):
    """Edits applied to the references before the reads are drawn."""
    nref = len(ro) - 1
    if kind == "rc_copies":
        # The reverse complement of a segment of ref 2i written into ref 2i+1:
        # a read of the segment scores the same on both strands.
        for r in range(0, nref - 1, 2):
            a, b = int(ro[r]), int(ro[r + 2])
            seg_len = int(rng.integers(800, 2000))
            src = int(ro[r]) + int(rng.integers(0, int(ro[r + 1] - ro[r]) - seg_len))
            dst = int(ro[r + 1]) + int(rng.integers(0, int(ro[r + 2] - ro[r + 1]) - seg_len))
            rb[dst:dst + seg_len] = COMP[rb[src:src + seg_len][::-1]]
            del a, b
        return
    if kind == "low_complexity":
        # Homopolymers and short tandem repeats: the off-diagonal head
        # comparison then runs to its boundary.
        for r in range(nref):
            a, b = int(ro[r]), int(ro[r + 1])
            for _ in range(25):
                at = a + int(rng.integers(0, b - a - 120))
                ln = int(rng.integers(20, 110))
                unit = rng.choice(list(b"ACGT"), int(rng.integers(1, 4))).astype(np.uint8)
                rb[at:at + ln] = np.resize(unit, ln)
        return
    if kind == "variant_copies":
        # Copies of a few segments with sparse substitutions, so that several
        # references share runs and chains of nearly the same length.
        segs = [rng.choice(list(b"ACGT"), int(rng.integers(600, 1500))).astype(np.uint8)
                for _ in range(4)]
        for r in range(nref):
            a, b = int(ro[r]), int(ro[r + 1])
            sg = segs[r % 4].copy()
            flips = rng.random(len(sg)) < 0.01 * (r % 5)
            sg[flips] = rng.choice(list(b"ACGT"), int(flips.sum()))
            at = a + int(rng.integers(0, b - a - len(sg)))
            rb[at:at + len(sg)] = sg
        return
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


def engineered(seed, T=30):
    """References built around chosen reads (not sampled), to reach the
    candidate filters.  Per read R:
      A: R with a substitution in its last probe window (run M-1, chain M-1),
      B: every probe k-mer of R, shuffled, between random spacers (run M,
         short chains),
      C: R with a 3-base insertion in the middle (run M, two chains),
      D: the first five probe windows of R, then the remaining probe k-mers
         off the diagonal (run M, chain 5);
    and per read R2:
      P: R2 up to the end of probe window 4 plus a tail's worth (one chain of 5),
      Q: R2 with one substitution inside window 4 (chains of 4 and 3)."""
    rng = np.random.default_rng(seed)
    s = T // 2
    k = T - s + 1
    rand = lambda n: rng.choice(list(b"ACGT"), n).astype(np.uint8)  # noqa: E731
    refs, reads = [], []
    for _ in range(30):
        R = rand(150)
        probes = [R[j:j + k] for j in range(0, 150 - T, s)]
        A = R.copy()
        last = (len(probes) - 1) * s
        A[last + k // 2] = {65: 67, 67: 71, 71: 84, 84: 65}[int(A[last + k // 2])]
        B = np.concatenate([np.concatenate([rand(int(rng.integers(20, 60))), probes[i]])
                            for i in rng.permutation(len(probes))])
        C = np.concatenate([R[:75], rand(3), R[75:]])
        D = np.concatenate([R[:4 * s + k]] + [np.concatenate([rand(int(rng.integers(25, 50))), p])
                                              for p in probes[5:]])
        # P: R up to the end of probe window 4 plus a tail's worth, then
        # random (one chain of 5); Q: R with a substitution inside window 4
        # only (chains of 4 and 3, which cover more of R than P does).
        R2 = rand(150)
        P = np.concatenate([R2[:4 * s + k + s - 1], rand(20)])
        Q = R2.copy()
        x4 = 4 * s + k // 2 + 1
        Q[x4] = {65: 67, 67: 71, 71: 84, 84: 65}[int(Q[x4])]
        reads.append(R2)
        for x in (A, B, C, D, P, Q):
            refs.append(np.concatenate([rand(int(rng.integers(200, 600))), x,
                                        rand(int(rng.integers(200, 600)))]))
        reads.append(R)
        reads.append(np.concatenate([R[:60], rand(1), R[61:]]))
    ro = np.zeros(len(refs) + 1, np.uint64)
    ro[1:] = np.cumsum([len(r) for r in refs])
    qo = np.zeros(len(reads) + 1, np.uint64)
    qo[1:] = np.cumsum([len(r) for r in reads])
    return np.concatenate(refs), ro, np.concatenate(reads), qo


def ties20(seed):
    """24 references of 1.5 kb: a 400-base segment copied unchanged into
    references 0-19 and a 300-base one into 20-23.  Reads: 60 from the first
    segment (20-way ties, so the reference sorts more than 16 scored
    candidates), 40 from the second, 50 from anywhere."""
    rng = np.random.default_rng(seed)
    rand = lambda n: rng.choice(list(b"ACGT"), n).astype(np.uint8)  # noqa: E731
    seg_a, seg_b = rand(400), rand(300)
    refs = []
    for r in range(24):
        x = rand(1500)
        seg = seg_a if r < 20 else seg_b
        at = 200 + 37 * (r % 7)
        x[at:at + len(seg)] = seg
        refs.append(x)
    ro = np.zeros(25, np.uint64)
    ro[1:] = np.cumsum([len(r) for r in refs])
    rb = np.concatenate(refs)
    reads = []
    for i in range(150):
        if i < 60:
            a = int(rng.integers(0, 250))
            reads.append(seg_a[a:a + 150].copy())
        elif i < 100:
            a = int(rng.integers(0, 150))
            reads.append(seg_b[a:a + 150].copy())
        else:
            r = int(rng.integers(0, 24))
            a = int(ro[r]) + int(rng.integers(0, 1350))
            reads.append(rb[a:a + 150].copy())
    qo = np.zeros(len(reads) + 1, np.uint64)
    qo[1:] = np.cumsum([len(r) for r in reads])
    return rb, ro, np.concatenate(reads), qo


def fasta_forms_text(prefix, by, off, rng, crlf):
    """FASTA text the reference reads in its several forms: headers with
    spaces and a '>' inside, line widths of 60, 61, 77 and unwrapped,
    lowercase stretches, CRLF or LF line ends, and no newline after the last
    line.  (The file must start with '>': Sequences.cpp:286-298.)"""
    nl = b"\r\n" if crlf else b"\n"
    out = bytearray()
    for i in range(len(off) - 1):
        seq = bytearray(by[int(off[i]):int(off[i + 1])].tobytes())
        for _ in range(3):
            a = int(rng.integers(0, len(seq)))
            b = min(len(seq), a + int(rng.integers(5, 80)))
            seq[a:b] = bytes(seq[a:b]).lower()
        out += b">%s%d sample %d > part" % (prefix.encode(), i, i % 3) + nl
        width = (60, 61, 77, 0)[i % 4] or len(seq)
        for j in range(0, len(seq), width):
            out += bytes(seq[j:j + width]) + nl
    return bytes(out[:-len(nl)])


def write_fasta(path, prefix, by, off):
    with open(path, "wb") as f:
        for i in range(len(off) - 1):
            f.write(b">%s%d\n" % (prefix.encode(), i))
            s = by[int(off[i]):int(off[i + 1])].tobytes()
            for j in range(0, len(s), 80):
                f.write(s[j:j + 80] + b"\n")


SHIPPED_DIR = "/root/reference/MirAge/data"


def ref_id_of(header, names):
    """Reference index of a header printed in results_Raw.log: the generated
    FASTA names references r<i>; `names` maps other names to indices."""
    name = header.lstrip(">").split()[0]
    return names[name] if names is not None else int(name[1:])


def run_reference(workdir, threshold, names=None):
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
    outdir = os.path.dirname(outs[0])
    mhm = [os.path.join(outdir, "hashmap_reference.mhm")]
    assert os.path.exists(mhm[0]), mhm
    logs = {}
    for kind in LOG_KINDS:
        with open(os.path.join(outdir, f"results_{kind}.log"), "rb") as f:
            logs[kind] = f.read().decode()
    res = {"_mhm": mhm[0], "_logs": logs}
    with open(outs[0]) as f:
        next(f)
        for line in f:
            line = line.rstrip("\n")
            if not line:
                continue
            cols = line.split(";")
            rid = int(cols[0])
            refs = sorted(ref_id_of(h, names) for h in cols[5].split(","))
            res[rid] = [cols[2], refs]
    return res


def main():
    only = set(sys.argv[1:])
    for name, T, rspec, qspec, seed, edits in datasets():
        if only and name not in only:
            continue
        names = None
        if edits == "shipped":
            from oracle import fasta as F
            texts = tuple(open(os.path.join(SHIPPED_DIR, fn), "rb").read()
                          for fn in ("reference.fasta", "query.fasta"))
            (rseq, rhead), (qseq, _) = F.parse(texts[0]), F.parse(texts[1])
            names = {h[1:-1].decode().split()[0]: i for i, h in enumerate(rhead)}
            rb = np.frombuffer(b"".join(rseq), np.uint8).copy()
            ro = np.zeros(len(rseq) + 1, np.uint64)
            ro[1:] = np.cumsum([len(x) for x in rseq])
            qb = np.frombuffer(b"".join(qseq), np.uint8).copy()
            qo = np.zeros(len(qseq) + 1, np.uint64)
            qo[1:] = np.cumsum([len(x) for x in qseq])
        elif edits == "engineered":
            rb, ro, qb, qo = engineered(seed, T)
        elif edits == "ties20":
            rb, ro, qb, qo = ties20(seed)
        else:
            rb, ro = D.make_refs(rspec, seed)
            rng = np.random.default_rng(seed)
            if edits != "n_runs":
                apply_edits(edits, rb, ro, None, None, rng)
            qb, qo, truth = D.make_reads(qspec, rb, ro, seed)
            if edits == "n_runs":
                apply_edits(edits, rb, ro, qb, qo, rng)
        base = os.path.join(RUNS, name)
        os.makedirs(base, exist_ok=True)
        if edits != "shipped":
            texts = None
        if edits == "shipped":
            for fn, t in zip(("reference.fasta", "query.fasta"), texts):
                with open(os.path.join(base, fn), "wb") as f:
                    f.write(t)
        elif edits == "fasta_forms":
            frng = np.random.default_rng(seed + 1)
            texts = (fasta_forms_text("r", rb, ro, frng, True),
                     fasta_forms_text("q", qb, qo, frng, False))
            for fn, t in zip(("reference.fasta", "query.fasta"), texts):
                with open(os.path.join(base, fn), "wb") as f:
                    f.write(t)
        else:
            write_fasta(os.path.join(base, "reference.fasta"), "r", rb, ro)
            write_fasta(os.path.join(base, "query.fasta"), "q", qb, qo)
        r1 = run_reference(os.path.join(base, "a"), T, names)
        r2 = run_reference(os.path.join(base, "b"), T, names)
        mhm = [open(r.pop("_mhm"), "rb").read() for r in (r1, r2)]
        logs = [r.pop("_logs") for r in (r1, r2)]
        if r1 != r2 or mhm[0] != mhm[1] or logs[0] != logs[1]:
            print(f"{name}: reference runs disagree, fixture not written")
            continue
        if name in MHM_FIXTURES:
            with gzip.GzipFile(os.path.join(OUT_DIR, f"ref_{name}.mhm.gz"), "wb", mtime=0) as f:
                f.write(mhm[0])
        fx = {
            "name": name,
            "threshold": T,
            "source": "oracle/_ref/mirage_ref (reference CPU build, ASan -O0), two agreeing runs",
            "refs": [rb[int(ro[i]):int(ro[i + 1])].tobytes().decode() for i in range(len(ro) - 1)],
            "reads": [qb[int(qo[i]):int(qo[i + 1])].tobytes().decode()
                      for i in range(len(qo) - 1)],
            "expected": {str(k): v for k, v in sorted(r1.items())},
            # the four result logs as the reference wrote them
            "logs": logs[0],
        }
        if texts is not None:  # the FASTA files as the reference read them
            fx["reference_fasta"] = texts[0].decode("latin-1")
            fx["query_fasta"] = texts[1].decode("latin-1")
        if edits == "shipped":
            # the sequences are those of the texts (tests/_data.py parses them)
            del fx["refs"], fx["reads"]
        with gzip.GzipFile(os.path.join(OUT_DIR, f"ref_{name}.json.gz"), "wb", mtime=0) as f:
            f.write(json.dumps(fx).encode())
        print(f"{name}: {len(r1)} classified of {len(qo) - 1}")


if __name__ == "__main__":
    main()
