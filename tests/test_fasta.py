"""FASTA reading: the oracle's restatement against what the reference read.

The reference stores the sequences and headers it read from reference.fasta
in its .mhm file; the fixture of t22_fasta_forms holds FASTA texts in the
forms it accepts (CRLF, mixed line widths and unwrapped lines, lowercase,
spaces and '>' inside headers, no final newline) and the reference's .mhm
and result logs for them.
"""
import pytest

from oracle import fasta as F
from oracle import mhm as M
from tests import _data


@pytest.mark.parametrize("name", _data.mhm_names())
def test_parse_equals_reference_read(name):
    fx, rb, ro, qb, qo = _data.load_golden(name)
    if "reference_fasta" in fx:
        text = fx["reference_fasta"].encode("latin-1")
    else:  # the generator's file: 80-column lines, LF
        text = F.write([(b"r%d" % i, r.encode()) for i, r in enumerate(fx["refs"])])
    seqs, heads = F.parse(text)
    m = M.read(_data.mhm_bytes(name), last_bucket_empty=False)
    assert seqs == m.refs and heads == m.headers


def test_forms_reads_and_headers():
    fx = _data.load_golden("t22_fasta_forms")[0]
    text = fx["query_fasta"].encode("latin-1")
    assert b"\r\n" not in text and not text.endswith(b"\n")
    assert b"\r\n" in fx["reference_fasta"].encode("latin-1")
    seqs, heads = F.parse(text)
    assert [s.decode() for s in seqs] == fx["reads"]
    printed = {int(ln.split(";")[0]): ln.split(";")[1] for ln in fx["logs"]["Raw"].splitlines()[1:]}
    assert printed and all(heads[i][:-1].decode() == h for i, h in printed.items())


@pytest.mark.parametrize("text,what", [
    (b">a\nACGT\n\nACGT\n", "empty line"),
    (b">a\nACGT\n\r\n", "empty line"),
    (b">a\nACGT\n@b\nACGT\n", "'@'"),
    (b">a\n>b\nACGT\n", "without sequence"),
    (b">a\nACGT\n>b\n", "without sequence"),
    (b">a\nACGT\n>b", "without sequence"),
    (b">a\nAC\0GT\n", "NUL"),
    (b"; note\n>a\nACGT\n", "start with"),
    (b"", "start with"),
])
def test_rejected_texts(text, what):
    with pytest.raises(F.FastaError) as e:
        F.parse(text)
    assert what in str(e.value)


def test_accepted_edge_forms():
    seqs, heads = F.parse(b">a x\r\nacgtn\r\nAC\r\n>b\nGGG")
    assert seqs == [b"ACGTNAC", b"GGG"] and heads == [b">a x\0", b">b\0"]
    assert F.parse(b">\xe9\nac{}\xe9\n") == ([b"AC[]\xe9"], [b">\xe9\0"])
