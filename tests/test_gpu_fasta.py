"""FASTA text parsed on the device (mh_fasta_parse) against the oracle's
restatement of the reference's reader, and the served path without host
preprocessing: FASTA bytes in HBM -> parse -> classify -> the reference's
result logs."""
import numpy as np
import pytest

from mirage_amd import datagen as D
from mirage_amd import native
from oracle import fasta as F
from tests import _data

pytestmark = pytest.mark.gpu


def to_device(text):
    """Text bytes in a device tensor (copied from a numpy array that owns them)."""
    import torch
    return torch.from_numpy(np.frombuffer(text, np.uint8).copy()).cuda()


def device_parse(handle, text):
    t = to_device(text)
    seq, so, hdr, ho = handle.parse_fasta(t)
    so, ho = so.cpu().numpy(), ho.cpu().numpy()
    sb, hb = seq.cpu().numpy().tobytes(), hdr.cpu().numpy().tobytes()
    return ([sb[so[i]:so[i + 1]] for i in range(len(so) - 1)],
            [hb[ho[i]:ho[i + 1]] for i in range(len(ho) - 1)])


def random_text(rng, n_rec, crlf, width):
    recs = []
    for i in range(n_rec):
        s = rng.choice(np.frombuffer(b"ACGTNacgtn", np.uint8), int(rng.integers(1, 3000)))
        recs.append((b"s%d %d>x\t%d" % (i, i * 7, i), s.tobytes()))
    t = F.write(recs, width=width, crlf=crlf)
    return t if rng.random() < 0.5 else t.rstrip(b"\r\n")


def test_parse_matches_oracle(handle):
    texts = []
    for name in _data.golden_names():
        fx = _data.load_golden(name)[0]
        if "reference_fasta" in fx:
            texts += [fx["reference_fasta"].encode("latin-1"), fx["query_fasta"].encode("latin-1")]
        else:
            texts.append(F.write([(b"q%d" % i, r.encode()) for i, r in enumerate(fx["reads"])]))
    rng = np.random.default_rng(3)
    for i in range(12):
        texts.append(random_text(rng, int(rng.integers(1, 400)), bool(i & 1),
                                 (60, 61, 1, 100000)[i % 4]))
    for i, t in enumerate(texts):
        print(f"text {i}: {len(t)} bytes", flush=True)
        assert device_parse(handle, t) == F.parse(t), i


def test_parse_at_size(handle):
    """Workload 2's million 150 bp reads as 61-column FASTA (180 MB)."""
    w = D.workload(2)
    rb, ro = D.make_refs(w.refs, w.seed)
    qb, qo, _ = D.make_reads(w.reads, rb, ro, w.seed)
    recs = [(b"read%d" % i, qb[int(qo[i]):int(qo[i + 1])].tobytes()) for i in range(len(qo) - 1)]
    text = F.write(recs, width=61)
    seqs, heads = device_parse(handle, text)
    assert len(seqs) == len(recs)
    assert b"".join(seqs) == qb.tobytes()
    assert heads[:3] == [b">read0\0", b">read1\0", b">read2\0"] and heads[-1] == b">read%d\0" % (
        len(recs) - 1)
    assert [len(s) for s in seqs] == np.diff(qo).tolist()


def test_parse_unaligned_and_resized(handle):
    """A text that does not start on a 16-byte boundary (byte loads), and
    texts whose records or headers outgrow the wrapper's first buffers (the
    exactly sized second call), parse as the oracle does."""
    import torch
    rng = np.random.default_rng(11)
    text = random_text(rng, 300, False, 61)
    src = to_device(text)
    big = torch.zeros(len(text) + 16, dtype=torch.uint8, device="cuda")
    for off in (1, 3, 8):
        view = big[off:off + len(text)]
        view.copy_(src)
        assert view.data_ptr() % 16 == off
        seq, so, hdr, ho = handle.parse_fasta(view)
        so, ho = so.cpu().numpy(), ho.cpu().numpy()
        sb, hb = seq.cpu().numpy().tobytes(), hdr.cpu().numpy().tobytes()
        got = ([sb[so[i]:so[i + 1]] for i in range(len(so) - 1)],
               [hb[ho[i]:ho[i + 1]] for i in range(len(ho) - 1)])
        assert got == F.parse(text), off
    tiny = b"".join(b">%d\nA\n" % i for i in range(5000))  # about 8 bytes a record
    assert device_parse(handle, tiny) == F.parse(tiny)
    wide = b">" + b"h" * 9000 + b"\nAC\n>" + b"g" * 7000 + b"\nT"  # headers fill the text
    assert device_parse(handle, wide) == F.parse(wide)
    exact = handle.parse_fasta(to_device(tiny), exact=True)
    assert exact[1].numel() == 5001


def test_parse_long_unwrapped_lines(handle):
    """Sequence lines of 1.5 Mb and 300 kb (unwrapped chromosomes): a line
    spans hundreds of wavefront windows; CRLF and lowercase included."""
    rng = np.random.default_rng(17)
    s1 = rng.choice(np.frombuffer(b"ACGTacgtN", np.uint8), 1_500_000).tobytes()
    s2 = rng.choice(np.frombuffer(b"ACGT", np.uint8), 300_000).tobytes()
    text = b">chr1 long\r\n" + s1 + b"\r\n>chr2\n" + s2[:150_000] + b"\n" + s2[150_000:]
    seqs, heads = device_parse(handle, text)
    exp = F.parse(text)
    assert heads == exp[1] and [len(x) for x in seqs] == [len(x) for x in exp[0]]
    assert seqs == exp[0]


@pytest.mark.parametrize("text,what", [
    (b">a\nACGT\n\nACGT\n", "empty line"),
    (b">a\nACGT\n@b\nACGT\n", "'@'"),
    (b">a\n>b\nACGT\n", "without a sequence"),
    (b">a\nACGT\n>b", "without a sequence"),
    (b">a\nAC\0GT\n", "NUL"),
    (b"; note\n>a\nACGT\n", "start with"),
])
def test_rejected_texts(handle, text, what):
    with pytest.raises(native.MirageError) as e:
        device_parse(handle, text)
    assert e.value.status == -1 and what in str(e.value)


@pytest.mark.parametrize("name", ["t22_fasta_forms", "t31_shipped_data"])
def test_fasta_to_logs_on_device(handle, tmp_path, name):
    """The reference and query FASTA of a dataset parsed on the device (the
    forms test, and the reference's own sample data at its shipped T = 31);
    the parsed reads classified where they lie; the logs written from the
    results and the parsed headers are the reference's."""
    import torch
    fx = _data.load_golden(name)[0]
    rt = to_device(fx["reference_fasta"].encode("latin-1"))
    qt = to_device(fx["query_fasta"].encode("latin-1"))
    rseq, rso, rhdr, rho = handle.parse_fasta(rt)
    qseq, qso, qhdr, qho = handle.parse_fasta(qt)
    ix = handle.build_index(rseq.cpu().numpy(), rso.cpu().numpy().view(np.uint64),
                            fx["threshold"])
    n = qso.numel() - 1
    d_out = torch.empty(n * native.READ_RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    top_cap = 64 * n + 1024
    d_top = torch.empty(top_cap * 4, dtype=torch.uint8, device="cuda")
    d_cands = torch.empty(64 * n * native.CAND_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    _, n_cands = ix.classify_device(qseq, qso, n, d_out, d_top, top_cap,
                                    flags=native.MH_CLASSIFY_BATCHED, d_cands=d_cands,
                                    cand_cap=64 * n)
    out = d_out.cpu().numpy().view(native.READ_RESULT_DTYPE)
    top = d_top.cpu().numpy().view(np.uint32)
    cands = d_cands.cpu().numpy().view(native.CAND_DTYPE)[:n_cands]
    hb, ho = qhdr.cpu().numpy().tobytes(), qho.cpu().numpy()
    rb_, rho_ = rhdr.cpu().numpy().tobytes(), rho.cpu().numpy()
    native.write_results(tmp_path, out, top, [hb[ho[i]:ho[i + 1]] for i in range(n)],
                         [rb_[rho_[i]:rho_[i + 1]] for i in range(len(rho_) - 1)], cands=cands)
    for k in ("Raw", "HighSensitivity", "HighPrecision", "Balanced"):
        assert (tmp_path / f"results_{k}.log").read_text(encoding="latin-1") == fx["logs"][k], k
    if name == "t31_shipped_data":
        assert (tmp_path / "results_Raw.log").read_text().split("\n")[1] == \
            "0;>RM2|S2|R14/1;0.627418;-1;2;>2,>1"
    ix.free()
