"""Pure-Python model of the reference classification path, for tiny inputs.

A second, independent transcription of MirAge/Analyser.cpp:1317-2490,
MirAge/MatrixCPU.cpp:18-370 and MirAge/MatrixResult.cpp:21-29, written loop
for loop against the reference sources.  It exists to be mutated: each entry
of MUTATIONS plants a plausible mistake, and the tests check that the golden
fixtures written by the reference binary reject every one of them.
"""
import ctypes
import ctypes.util
import struct

_libm = ctypes.CDLL(ctypes.util.find_library("m"))
_libm.logf.restype = ctypes.c_float
_libm.logf.argtypes = [ctypes.c_float]


def f32(x):
    return struct.unpack("<f", struct.pack("<f", x))[0]


def hash28(b):
    h = 2166136261
    for c in b:
        h = ((h * 16777619) & 0xFFFFFFFF) ^ c
    return ((h >> 28) ^ h) & 0x0FFFFFFF


COMP = {ord(a): ord(b) for a, b in zip("ATUGCYRSWKMBDHVN-~", "TAACGRYSWMKVHDBN-~")}

MUTATIONS = ("head_on_diagonal", "head_counts_boundary", "cut_is_max_run", "istop_uses_k",
             "tail_unbounded", "chains_keep_all_refs", "no_nnn_cutoff", "rc_ties_go_reverse")


class Model:
    def __init__(self, refs, T, mutation=None, last_bucket_empty=True):
        self.T, self.s = T, T // 2
        self.k = T - self.s + 1
        self.mut = mutation
        self.refs = [bytes(r).upper() for r in refs]
        k = self.k
        nnn = hash28(b"N" * k)
        buckets = {}
        phantom = 0
        for rid, r in enumerate(self.refs):
            n = len(r)
            if n <= k:
                continue
            for i in range(n - k):
                h = hash28(r[i:i + k])
                if h == nnn and mutation != "no_nnn_cutoff":
                    # the hashed pointer stops advancing (Hashmap.cpp:2882-2895);
                    # the windows left still count in the converted size
                    # (:1233-1239, :1252) as value-initialised (0, 0) entries
                    phantom += (n - k) - i
                    break
                if h == nnn:
                    continue
                buckets.setdefault(h, []).append((rid, i))
        if phantom:
            buckets.setdefault(0, []).extend([(0, 0)] * phantom)
        if T < 20 and buckets:
            sizes = sorted(len(v) for v in buckets.values())
            thr = min(20, sizes[int((len(sizes) / 100.0) * 30)])
            buckets = {h: v for h, v in buckets.items() if len(v) <= thr}
        if last_bucket_empty and buckets:
            buckets[max(buckets)] = []
        self.buckets = buckets

    # Analyser::TPrepareGPUAnalysis for one query
    def seeds(self, q):
        T, s, k = self.T, self.s, self.k
        m = len(q)
        if m <= T:
            return [], {}
        istop = m - (k if self.mut == "istop_uses_k" else T)
        hits = []
        for j in range(0, istop, s):
            h = hash28(q[j:j + k])
            if h in self.buckets:
                hits.append((j, self.buckets[h]))
        if not hits or not any(b for _, b in hits):
            return [], {}
        occ, occmax, last = {}, {}, {}
        highest = 0
        for i, (_, b) in enumerate(hits):
            for r, _ in b:
                if i == 0:
                    occ[r], occmax[r], last[r] = 1, 1, 0
                    continue
                if last.get(r, -1) == i:
                    continue
                if last.get(r, -1) == i - 1:
                    occ[r] += 1
                    occmax[r] = max(occmax[r], occ[r])
                    highest = max(highest, occmax[r])
                else:
                    occ[r] = 1
                    occmax[r] = max(occmax.get(r, 0), 1)
                last[r] = i
        cut = max(1, highest - 1 if highest else 0)
        if self.mut == "cut_is_max_run":
            cut = max(1, highest)
        pairs = sorted((r, p, j) for j, b in hits for r, p in b if occmax[r] >= cut)
        cands = sorted(set(r for r, _, _ in pairs))
        chains = {}
        for r in cands:
            el = [(p, j) for rr, p, j in pairs if rr == r]
            ch = [[el[0]]]
            for a, b in zip(el, el[1:]):
                if a[0] == (b[0] - s) & 0xFFFFFFFF and a[1] == (b[1] - s) & 0xFFFFFFFF:
                    ch[-1].append(b)
                else:
                    ch.append([b])
            chains[r] = ch
        G = max(max(len(c) for c in ch) for ch in chains.values())
        if G < 2:
            return cands, {}
        out = {}
        for r in cands:
            if max(len(c) for c in chains[r]) != G and self.mut != "chains_keep_all_refs":
                continue
            out[r] = [(c[0][0], c[0][1], c[-1][0] + s, c[-1][1] + s) for c in chains[r]]
        return cands, out

    # MatrixCPU::CalculateScore
    def score(self, q, r, seeds):
        T = self.T
        ref = self.refs[r]
        m, n = len(q), len(ref)
        row = [0] * m
        sr = sc = 0
        for sref, sq, lref, lq in seeds:
            rs, re = sq, lq
            if rs > re:
                continue
            mid = re - rs
            t = 0
            tmax = min(T, n - lref) if self.mut != "tail_unbounded" else n
            while re + t < m and lref + t < n and t < tmax and q[re + t] == ref[lref + t]:
                t += 1
            h = 0
            if rs > 0:
                hs = min(T, sref)
                if hs > 0:
                    l = 0
                    while True:
                        rp = sref - l - (1 if self.mut == "head_on_diagonal" else 0)
                        if rp < 0 or q[rs - 1 - l] != ref[rp]:
                            h = l
                            break
                        if rs - 1 - l == 0 or sref - l == 0 or l == hs - 1:
                            h = l + (1 if self.mut == "head_counts_boundary" else 0)
                            break
                        l += 1
            tot = t + h + mid
            if tot >= T:
                sc += tot
                for x in range(rs - h, min(m, rs - h + tot)):
                    row[x] = 1
                sr += tot * (tot + 1) // 2
        if sc == 0:
            return -2.0, sr, sc, 0
        H = sum(row)
        sf = f32(f32(float(sr)) / f32(float(sc)))
        pct = f32((100.0 / f32(float(m))) * f32(float(H)))
        return f32((sf / 100.0) * pct), sr, sc, H

    def rel(self, m, n, score):
        L = f32(min(f32(float(m)), f32(float(n))))
        smax = f32(((L * (L + 1.0)) / 2.0) / L + 1)
        x = f32(score + 1.0)
        return f32(_libm.logf(x) / _libm.logf(smax))

    def classify(self, read):
        read = bytes(read).upper()
        rc = bytes(COMP[c] for c in reversed(read))
        res = []
        for q in (read, rc):
            _, sd = self.seeds(q)
            best, rels = -1.0, {}
            for r in sorted(sd):
                sc = self.score(q, r, sd[r])[0]
                if sc > -1:
                    rels[r] = self.rel(len(q), len(self.refs[r]), sc)
                    best = max(best, rels[r])
            res.append((best, rels))
        f, b = res
        take_rc = f[0] < b[0] or (self.mut == "rc_ties_go_reverse" and f[0] == b[0])
        best, rels = b if take_rc else f
        if not rels or best == 0:
            return None
        return ("%f" % best, tuple(sorted(r for r, v in rels.items() if v == best)))
