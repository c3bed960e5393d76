"""MirAge's FASTA reader, restated in Python (test infrastructure).

Sequences::ReadFastaSequencesFromMem (MirAge/Sequences.cpp:325-562) finds
header positions (GetNumberOfHeaders, :283-320: a '>' after a newline; a
text that does not start with '>' is refused with "0 Headers found",
:286-298), hands runs of headers to threads, and each thread
(TReadFastaSequencesFromMem, :57-211) reads its text line by line:

* a line starting with '>' (or '@') is a header: it is kept with its '>' and
  a NUL in place of a final '\\r', or a NUL appended (:117-127); the sequence
  gathered so far is closed (uppercased, :40-52);
* any other line is sequence: its bytes without a final '\\r' are appended
  (:143-151);
* a last line without a newline is read unless it is a header (:166-181).

That is restated here for the files the library accepts: every header line
followed by at least one non-empty sequence line, no empty lines, no line
starting with '@' and no NUL byte.  Outside that set the reference reads
past its buffers (an empty line reads line[-1], :102-118) or pairs headers
and sequences differently depending on how its threads split the file
(:356-380), so there is nothing defined to match; parse() raises instead,
as the library returns MH_ERR_ARG.
"""


class FastaError(ValueError):
    pass


def upper(b):
    """SequenceToUpperCase (Sequences.cpp:40-52): char > 96 loses 32; bytes
    of 128 and above are negative chars and stay."""
    return bytes(c - 32 if 96 < c < 128 else c for c in b)


def parse(text):
    """(sequences, headers) of a FASTA text, as bytes; headers as MirAge
    keeps them ('>' name NUL)."""
    if b"\0" in text:
        raise FastaError("NUL byte")
    if not text.startswith(b">"):  # 0 headers found, the program stops (:286-298)
        raise FastaError("text does not start with '>'")
    lines = text.split(b"\n")
    last_has_newline = text.endswith(b"\n")
    if last_has_newline:
        lines = lines[:-1]
    seqs, heads = [], []
    for i, ln in enumerate(lines):
        body = ln[:-1] if ln.endswith(b"\r") else ln
        if ln[:1] == b"@":
            raise FastaError("line starting with '@'")
        if not body:
            raise FastaError("empty line")
        if ln[:1] == b">":
            if i == len(lines) - 1 and not last_has_newline:
                raise FastaError("header without sequence")
            heads.append(body + b"\0")
            seqs.append(b"")
        else:
            seqs[-1] += body
    if any(not s for s in seqs):
        raise FastaError("header without sequence")
    return [upper(s) for s in seqs], heads


def write(records, width=80, crlf=False):
    """FASTA text of (name, sequence) records, sequence lines of `width`."""
    nl = b"\r\n" if crlf else b"\n"
    out = bytearray()
    for name, seq in records:
        out += b">" + name + nl
        for j in range(0, len(seq), width):
            out += seq[j:j + width] + nl
    return bytes(out)
