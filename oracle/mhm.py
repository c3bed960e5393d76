"""CPU restatement of MirAge's `.mhm` v4 hashmap file (test infrastructure).

This module is the oracle for the `.mhm` reader and writer of libmirage_hip:
only tests, the smoke check and the bench's CPU leg use it.  Plain Python,
for small files.

Layout (integers little-endian; HB = the file's hash width, 4 in a 32-bit
build or 8; SB = the stream-position width):

  header   73-byte indicator string (Hashmap.cpp:408-414), u64 number of
           TOC entries (:428-430), i32 version = 4 (:433-434), i32 HB
           (:444), i32 threshold (:456), then four single bytes: SB,
           bytes_numRefIDsPerHash, writeSequenceBitsSize,
           writeHeaderBitsSize (:475-478)
  TOC      per bucket in ascending hash order: hash [HB], stream position of
           its record [SB] (writer Hashmap.cpp:2062-2091, reader :519-535)
  taxids   count [HB], count x i32 (writer :2097-2117, reader :552-558)
  sequence TOC  position after this TOC [SB], count [HB], count sizes [HB
           each], count stream positions [SB each] (writer :2122-2172,
           reader LoadReferenceDataFromHashmapDB_SuperLazy :629-660)
  header TOC    the same for the reference headers (writer :2176-2222)
  data     the sequences, then the headers (writer :2226-2240)
  records  one per TOC entry (writer :2262-2516, reader Analyser.cpp:1614-1720):
           4 width bytes bPO, bID, bPOS, bNS; refIDMin [bPO]; refPOSMin [bPO];
           n [bNS]; n ids - refIDMin [bID each]; bRep (1 byte); number of
           position runs [bPOS]; per run: count [bRep], pos - refPOSMin [bPOS].
           Ids pair index-wise with the run-expanded positions.

The writer sorts each bucket by position before encoding
(SortAndCorrectPairVector, Hashmap.cpp:1561-1573, called at :1998 and
:2005), so refPOSMin is the first position and every stored position is
exact.  Widths are ((bits(v) - 1) // 8) + 1 bytes (GetSaveHashmapPreCalculates,
:1620-1697).

The loader derives each record's size from the next TOC position and never
sets it for the last entry (Hashmap.cpp:525-538), so the highest-hash bucket
is present but decodes empty (Analyser.cpp:1608-1609 skips size 0):
`last_bucket_empty=True` keeps that behaviour.
"""
import struct

INDICATOR = b"MirAge - Maarten Pater - www.mirdesign.nl - mirage@mirdesign.nl - Hashmap"
VERSION = 4


class MhmError(ValueError):
    pass


def nbytes(v):
    """Bytes the writer spends on value v: ((countBits(v) - 1) / 8) + 1,
    with countBits(0) = 0 giving 1 (Hashmap.cpp:1638-1650)."""
    return max(1, (int(v).bit_length() + 7) // 8)


class Mhm:
    """A parsed `.mhm` file.  buckets: hash -> list of (ref, pos) in file order."""

    def __init__(self):
        self.threshold = 0
        self.hash_bytes = 4
        self.sp_bytes = 4
        self.ids_per_hash_bytes = 1
        self.seq_size_bytes = 4
        self.header_size_bytes = 4
        self.hashes = []       # TOC order (ascending)
        self.streampos = []
        self.taxids = []
        self.refs = []         # bytes per reference, as stored
        self.headers = []      # bytes per reference header
        self.buckets = {}


def _u(b, o, n):
    if o + n > len(b):
        raise MhmError(f"truncated file: {n} bytes at offset {o}")
    return int.from_bytes(b[o:o + n], "little"), o + n


def decode_record(b, o, end):
    """Decodes one bucket record at [o, end) (Analyser.cpp:1614-1720)."""
    def take(n):
        nonlocal o
        if n < 1 or n > 8:
            raise MhmError(f"field width {n} at offset {o}")
        if o + n > end:
            raise MhmError(f"record overruns its extent at offset {o}")
        v = int.from_bytes(b[o:o + n], "little")
        o += n
        return v
    if o + 4 > end:
        raise MhmError(f"record overruns its extent at offset {o}")
    bpo, bid, bpos, bns = b[o], b[o + 1], b[o + 2], b[o + 3]
    o += 4
    id_min = take(bpo)
    pos_min = take(bpo)
    n = take(bns)
    ids = [take(bid) + id_min for _ in range(n)]
    brep = take(1)
    nrun = take(bpos)
    pos = []
    for _ in range(nrun):
        cnt = take(brep)
        p = take(bpos) + pos_min
        pos.extend([p] * cnt)
    if len(pos) != n:
        raise MhmError(f"record at offset {o}: {n} ids but {len(pos)} positions")
    return list(zip(ids, pos)), o


def read(data, last_bucket_empty=True):
    """Parses a whole `.mhm` v4 file given as bytes."""
    b = bytes(data)
    m = Mhm()
    if b[:len(INDICATOR)] != INDICATOR:
        raise MhmError("not a MirAge hashmap file (indicator string)")
    o = len(INDICATOR)
    nh, o = _u(b, o, 8)
    ver, hb, thr = struct.unpack_from("<iii", b, o)
    o += 12
    if ver != VERSION:
        raise MhmError(f"unsupported hashmap version {ver}")
    if hb not in (4, 8):
        raise MhmError(f"unsupported hash width {hb}")
    m.threshold, m.hash_bytes = thr, hb
    sb, m.ids_per_hash_bytes, m.seq_size_bytes, m.header_size_bytes = b[o:o + 4]
    o += 4
    if not 1 <= sb <= 8:
        raise MhmError(f"stream position width {sb}")
    m.sp_bytes = sb
    for _ in range(nh):
        h, o = _u(b, o, hb)
        p, o = _u(b, o, sb)
        m.hashes.append(h)
        m.streampos.append(p)
    ntax, o = _u(b, o, hb)
    m.taxids = list(struct.unpack_from(f"<{ntax}i", b, o))
    o += 4 * ntax
    sizes = {}
    for what in ("seq", "hdr"):
        _after, o = _u(b, o, sb)
        n, o = _u(b, o, hb)
        sz = []
        for _ in range(n):
            v, o = _u(b, o, hb)
            sz.append(v)
        sp = []
        for _ in range(n):
            v, o = _u(b, o, sb)
            sp.append(v)
        sizes[what] = (sz, sp)
    for what, dest in (("seq", m.refs), ("hdr", m.headers)):
        for sz, sp in zip(*sizes[what]):
            if sp + sz > len(b):
                raise MhmError("sequence data past the end of the file")
            dest.append(b[sp:sp + sz])
    for i, (h, sp) in enumerate(zip(m.hashes, m.streampos)):
        end = m.streampos[i + 1] if i + 1 < nh else len(b)
        if i + 1 == nh and last_bucket_empty:
            m.buckets[h] = []
            continue
        m.buckets[h], _ = decode_record(b, sp, end)
    return m


def encode_record(entries):
    """Record bytes of one bucket, entries (ref, pos) already in the order
    the writer keeps (sorted by position)."""
    ids = [r for r, _ in entries]
    pos = [p for _, p in entries]
    id_min, id_max = min(ids), max(ids)
    pos_min, pos_max = pos[0], pos[-1]
    n = len(entries)
    bid = nbytes(id_max - id_min)
    bpos = nbytes(pos_max - pos_min)
    bpo = nbytes(max(id_min, pos_min))
    bns = nbytes(n)
    brep = nbytes(n)
    runs = []
    for p in pos:
        if runs and runs[-1][1] == p - pos_min:
            runs[-1][0] += 1
        else:
            runs.append([1, p - pos_min])
    le = lambda v, w: (v & ((1 << (8 * w)) - 1)).to_bytes(w, "little")  # noqa: E731
    out = bytearray([bpo, bid, bpos, bns])
    out += le(id_min, bpo) + le(pos_min, bpo) + le(n, bns)
    for r in ids:
        out += le(r - id_min, bid)
    out += bytes([brep]) + le(len(runs), bpos)
    for c, p in runs:
        out += le(c, brep) + le(p, bpos)
    return bytes(out)


def write(m):
    """Bytes of a `.mhm` v4 file (Hashmap.cpp:1835-2540) from m.hashes (ascending),
    m.buckets (entries in record order), m.refs, m.headers and m.taxids."""
    hb, sb = m.hash_bytes, m.sp_bytes
    n_refs = len(m.refs)
    le = lambda v, w: (v & ((1 << (8 * w)) - 1)).to_bytes(w, "little")  # noqa: E731
    head = bytearray(INDICATOR)
    head += struct.pack("<Qiii", len(m.hashes), VERSION, hb, m.threshold)
    # widths of the largest bucket, sequence and header (Hashmap.cpp:1754-1755,
    # :1821-1830)
    head += bytes([sb, nbytes(max((len(v) for v in m.buckets.values()), default=0)),
                   nbytes(max((len(r) for r in m.refs), default=0)),
                   nbytes(max((len(h) for h in m.headers), default=0))])
    toc_bytes = len(m.hashes) * (hb + sb)
    meta = hb + 4 * n_refs + 2 * (sb + hb + n_refs * (hb + sb))
    data0 = len(head) + toc_bytes + meta
    seq_total = sum(len(r) for r in m.refs)
    hdr_total = sum(len(h) for h in m.headers)
    rec0 = data0 + seq_total + hdr_total
    recs = [encode_record(m.buckets[h]) for h in m.hashes]
    out = head
    sp = rec0
    for h, r in zip(m.hashes, recs):
        out += le(h, hb) + le(sp, sb)
        sp += len(r)
    out += le(n_refs, hb) + struct.pack(f"<{n_refs}i", *m.taxids)
    p = data0
    for items in (m.refs, m.headers):
        after = len(out) + sb + hb + n_refs * (hb + sb)
        out += le(after, sb) + le(n_refs, hb)
        for x in items:
            out += le(len(x), hb)
        for x in items:
            out += le(p, sb)
            p += len(x)
    for x in m.refs:
        out += x
    for x in m.headers:
        out += x
    assert len(out) == rec0
    for r in recs:
        out += r
    return bytes(out)
