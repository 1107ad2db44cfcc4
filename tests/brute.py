"""Brute-force definitions used to pin the oracle on tiny texts (test infrastructure).

Everything here is computed from first principles on an explicit suffix list, never from the
FM-index: the suffix array of T$ (T = forward + reverse complement, '$' smallest), the full BWT
with '$', Occ(c, k) as a plain count, and the interval of a pattern as (number of suffixes smaller
than P, same for revcomp(P), number of suffixes starting with P).
"""
from __future__ import annotations

import bisect

import numpy as np

ALPHA = "0123"          # nt4 codes as characters; '$' (0x24) sorts before '0'


def revcomp(p: str) -> str:
    return "".join("3210"[int(c)] for c in reversed(p))


class TinyText:
    def __init__(self, ref: np.ndarray):
        fwd = "".join(ALPHA[int(c)] for c in ref)
        self.T = fwd + revcomp(fwd)
        self.n = len(self.T)
        suf = sorted(range(self.n + 1), key=lambda i: self.T[i:] + "$")
        self.sa = suf                                        # row -> text position (row 0 = '$')
        self.sufs = [self.T[i:] + "$" for i in suf]
        self.L = "".join("$" if i == 0 else self.T[i - 1] for i in suf)   # BWT with '$'
        self.primary = self.L.index("$")
        self.B = self.L.replace("$", "")                                  # BWT without '$'

    def occ(self, k: int) -> list:
        """Occ(c, k) = number of c in L[0:k] for c = 0..3."""
        s = self.L[:k]
        return [s.count(c) for c in ALPHA]

    def lexrank(self, p: str) -> int:
        return bisect.bisect_left(self.sufs, p)

    def count(self, p: str) -> int:
        return bisect.bisect_left(self.sufs, p + "~") - bisect.bisect_left(self.sufs, p)

    def interval(self, p: str):
        """(x0, x1, size) of pattern p in the bidirectional index."""
        return self.lexrank(p), self.lexrank(revcomp(p)), self.count(p)


def seq_str(q: np.ndarray) -> str:
    return "".join("N" if c > 3 else ALPHA[int(c)] for c in q)


def quirk_prefix(tt: TinyText, p: str) -> bool:
    """True when the interval of p is the single row `primary` (p occurs once, at text position 0):
    extending it backward by one base takes the l-k==1 shortcut of mb_bwt_rank2a at k == primary
    (bwt.c:205-209), which counts the $-removed BWT symbol at index primary although row primary
    holds '$'.  The FM-index then reports one occurrence of a pattern that does not occur."""
    return tt.count(p) == 1 and tt.T.startswith(p)


def smem_from(tt: TinyText, q: str, qlen: int, x: int, min_len: int, min_occ: int, flags=None):
    """The (min_len, min_occ)-SMEM search from position x as defined by the reference's serial
    algorithm (bwt.c:253-305), with every occurrence count taken from the suffix list.  Returns
    (next x, (x0, x1, size, qs, qe) or None).  Extensions that start from a quirk_prefix interval
    (backward) or its reverse-complement counterpart (forward) are noted in flags["quirk"]."""
    if flags is None:
        flags = {}
    if qlen - x < min_len:
        return qlen, None
    xn = -1
    for i in range(x, x + min_len):
        if q[i] == "N":
            xn = i
    if xn >= 0:
        return xn + 1, None

    def back(st, pos):
        # longest left extension of q[.. pos] keeping >= min_occ occurrences; the first base alone
        # is taken unconditionally (mb_bwt_set_intv)
        i = pos - 1
        while i >= st:
            if q[i] != "N" and quirk_prefix(tt, q[i + 1:pos + 1]):
                flags["quirk"] = True
            if q[i] == "N" or tt.count(q[i:pos + 1]) < min_occ:
                break
            i -= 1
        return i

    i = back(x, x + min_len - 1)
    if i >= x:
        return i + 1, None
    j = x + min_len
    while j < qlen:
        if q[j] != "N" and quirk_prefix(tt, revcomp(q[x:j])):
            flags["quirk"] = True
        if q[j] == "N" or tt.count(q[x:j + 1]) < min_occ:
            break
        j += 1
    x0, x1, size = tt.interval(q[x:j])
    hit = (x0, x1, size, x, j)
    if j == qlen:
        return qlen, hit
    i = j if q[j] == "N" else back(x + 1, j)
    return i + 1, hit


def seed_intv(tt: TinyText, q: str, min_len: int, max_sub_occ: int, flags=None):
    """Two-pass seeding (seed.c:24-54) on the brute-force SMEM search: list of (x0,x1,size,qs,qe)."""
    out, x, qlen = [], 0, len(q)
    while True:
        x, h = smem_from(tt, q, qlen, x, min_len, 1, flags)
        if h is not None and h[2] > 0:
            out.append(h)
        if x >= qlen:
            break
    n_a0 = len(out)
    for k in range(n_a0):
        _, _, size, st, en = out[k]
        if en - st < min_len * 2 or size > max_sub_occ:
            continue
        x = st
        sub = max((en - st) // 2, min_len)
        while True:
            x, h = smem_from(tt, q, en, x, sub, size + 1, flags)
            if h is not None and h[2] > size:
                out.append(h)
            if x >= en:
                break
    return out


def gap_cost(k: int, q: int, e: int, q2: int, e2: int) -> int:
    return min(q + k * e, q2 + k * e2)


def gotoh(query, target, mat, q, e, q2, e2, m=5):
    """Textbook global DP with gap cost min(q+k*e, q2+k*e2) (two affine gap states per direction).
    Returns the (qlen x tlen) score matrix H with H[i][j] = best score of query[..i] vs target[..j],
    starting from the empty prefix pair at score 0 and penalising leading gaps."""
    NEG = -(1 << 30)
    ql, tl = len(query), len(target)
    H = [[NEG] * (tl + 1) for _ in range(ql + 1)]
    E = [[[NEG] * (tl + 1) for _ in range(ql + 1)] for _ in range(2)]   # gap in query (consumes target)
    F = [[[NEG] * (tl + 1) for _ in range(ql + 1)] for _ in range(2)]   # gap in target (consumes query)
    go = ((q, e), (q2, e2))
    H[0][0] = 0
    for i in range(ql + 1):
        for j in range(tl + 1):
            if i == 0 and j == 0:
                continue
            best = NEG
            for s, (o, x) in enumerate(go):
                if j > 0:
                    E[s][i][j] = max(H[i][j - 1] - o - x, E[s][i][j - 1] - x)
                    best = max(best, E[s][i][j])
                if i > 0:
                    F[s][i][j] = max(H[i - 1][j] - o - x, F[s][i - 1][j] - x)
                    best = max(best, F[s][i][j])
            if i > 0 and j > 0:
                best = max(best, H[i - 1][j - 1] + int(mat[int(target[j - 1]) * m + int(query[i - 1])]))
            H[i][j] = best
    return [row[1:] for row in H[1:]]


def cigar_score(cigar, query, target, mat, q, e, q2, e2, m=5):
    """Score of the alignment a CIGAR (len<<4 | op; 0=M, 1=I, 2=D) describes, and the query/target
    lengths it consumes."""
    i = j = sc = 0
    for c in cigar:
        ln, op = int(c) >> 4, int(c) & 0xf
        if op == 0:
            for k in range(ln):
                sc += int(mat[int(target[j + k]) * m + int(query[i + k])])
            i += ln
            j += ln
        elif op == 1:
            sc -= gap_cost(ln, q, e, q2, e2)
            i += ln
        elif op == 2:
            sc -= gap_cost(ln, q, e, q2, e2)
            j += ln
        else:
            raise ValueError(f"unexpected CIGAR op {op}")
    return sc, i, j


def sw_local_rows(query, target, mat, gapo, gape, n_pos, id_adjacent, m=5):
    """Textbook local Smith-Waterman with affine gaps (a gap of k bases costs gapo + k*gape) over
    the query padded to n_pos positions, each padded position scoring -1 against every base.
    id_adjacent: whether a deletion (E, consuming target) may open right after an insertion (F,
    consuming query).  Returns the H matrix as a list of rows, one per target base."""
    ql = len(query)
    sc = [[(int(mat[t * m + int(query[k])]) if k < ql else -1) for k in range(n_pos)] for t in range(m)]
    Hp = [0] * n_pos
    E = [0] * n_pos
    rows = []
    for i in range(len(target)):
        s = sc[min(int(target[i]), m - 1)]
        H = [0] * n_pos
        f = 0
        for k in range(n_pos):
            d = (Hp[k - 1] if k else 0) + s[k]
            hn = max(0, d, E[k])                # paths not ending in an insertion
            h = max(hn, f)
            H[k] = h
            E[k] = max(0, E[k] - gape, (h if id_adjacent else hn) - gapo - gape)
            f = max(0, f - gape, h - gapo - gape)
        rows.append(H)
        Hp = H
    return rows


def ungap_votes(query, target, mt=0, k=7):
    """mb_ungap's vote by its definition: every target position j ending a k-mer without N votes for
    diagonal j - i, where i is the last query position ending the same (methylation-converted) k-mer,
    when i <= j.  Returns (votes per diagonal as a dict, number of distinct query k-mers)."""
    conv = {0: (0, 1, 2, 3), 1: (0, 3, 2, 3), 2: (0, 1, 0, 3)}[mt]

    def kmers(s):
        out = {}
        for j in range(k - 1, len(s)):
            w = s[j - k + 1:j + 1]
            if all(int(c) < 4 for c in w):
                out[j] = tuple(conv[int(c)] for c in w)
        return out

    last = {}
    for i, km in sorted(kmers(query).items()):
        last[km] = i
    votes = {}
    for j, km in kmers(target).items():
        i = last.get(km)
        if i is not None and j >= i:
            votes[j - i] = votes.get(j - i, 0) + 1
    return votes, len(last)


def random_alignment(rng, n_ops=None, lowcomp=False):
    """A CIGAR of M/I/D runs (adjacent gaps and zero-length runs included) and query/target bases
    that it aligns, with substitutions and Ns in the M runs: (cigar, qseq, tseq)."""
    alpha = rng.integers(0, 4, 2) if lowcomp else None
    ops = []
    for _ in range(n_ops or int(rng.integers(1, 9))):
        op = int(rng.choice([0, 0, 1, 2]))
        ln = int(rng.integers(0 if rng.random() < 0.05 else 1, 12))
        ops.append((op, ln))
    q, t = [], []

    def base():
        return int(rng.choice(alpha)) if lowcomp else int(rng.integers(0, 4))
    for op, ln in ops:
        for _ in range(ln):
            if op == 0:
                b = base()
                q.append(b)
                r = rng.random()
                t.append(int(rng.integers(0, 4)) if r < 0.05 else 4 if r < 0.07 else b)
            elif op == 1:
                q.append(base())
            else:
                t.append(base())
    return [ln << 4 | op for op, ln in ops], np.array(q, np.uint8), np.array(t, np.uint8)
