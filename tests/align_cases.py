"""Seeded mb_align_pair calls (align.c:329-390) that reach every branch: the ungapped extension fast
path (taken, refused by the mismatch count, refused by the end bonus), the ungapped global fast path
(few mismatches, score above the best gapped score, refused), the band cap from the mismatch count
for sides under 100 bases, the max_sw_mat skip, the generic-scoring switch (transition mismatch,
methylation) and the extz2/extd2 dispatch."""
import numpy as np

from oracle import oracle as O

# (p = a, b, b_ts, q, e, q2, e2, meth; matrix conversion mt)
OPTS = [
    ((2, 8, 0, 12, 2, 24, 1, 0), 0),      # short-read preset gaps, extd2
    ((2, 8, 0, 4, 2, 24, 1, 0), 0),       # the mb_opt_init defaults (options.c:22-25)
    ((2, 8, 0, 6, 2, 6, 2, 0), 0),        # single affine: extz2
    ((2, 8, 4, 12, 2, 24, 1, 0), 0),      # transition mismatch: generic scoring
    ((2, 8, 0, 12, 2, 24, 1, 1), 1),      # methylation (C->T): generic scoring
    ((1, 4, 0, 6, 1, 26, 1, 0), 0),       # another scale
]


def _mutate(rng, s, sub, indel):
    out = []
    for c in s:
        r = rng.random()
        if r < indel / 2:
            continue
        if r < indel:
            out.append(int(rng.integers(4)))
        out.append(int((c + rng.integers(1, 4)) % 4) if rng.random() < sub else int(c))
    return np.array(out, np.uint8)


def cases(seed, n):
    """n calls: (p, max_sw_mat, query, target, mat, w, end_bonus, zdrop, flag)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        p, mt = OPTS[i % len(OPTS)]
        a, b, b_ts = p[0], p[1], p[2]
        mat = O.gen_nt4_mat(a, b, b_ts, 1, mt)
        kind = rng.integers(7)
        qlen = int(rng.choice([int(rng.integers(0, 12)), int(rng.integers(12, 100)), int(rng.integers(100, 400))]))
        q = rng.integers(0, 4, qlen).astype(np.uint8)
        if rng.random() < 0.15 and qlen > 0:
            q[rng.integers(qlen)] = 4                                   # an ambiguous base
        w = int(rng.choice([5, 20, 61, 101, 151, 500]))
        eb = int(rng.choice([0, 5, 10, 30]))
        zd = int(rng.choice([-1, 20, 100, 200]))
        msm = 0
        if kind <= 2:      # extension: tlen >= qlen with a few or many mismatches; left or right
            nsub = int(rng.choice([0, 1, 2, 3, 6])) if qlen > 0 else 0
            t = q.copy()
            for k in rng.choice(max(qlen, 1), size=min(nsub, qlen), replace=False) if qlen > 0 else []:
                t[k] = (t[k] + 1) % 4
            if kind == 2 and qlen > 4:                                  # mismatches crowded at the end
                t[-3:] = (t[-3:] + 1) % 4
            t = np.concatenate([t, rng.integers(0, 4, int(rng.integers(0, 60))).astype(np.uint8)])
            flag = int(rng.choice([0x40, 0x40 | 0x02 | 0x80, 0x40 | 0x10]))
        elif kind == 3:    # extension with the target shorter than the query: no fast path, no cap
            t = _mutate(rng, q, 0.02, 0.01)[: max(0, qlen - int(rng.integers(1, 20)))]
            flag = 0x40
        elif kind == 4:    # global, equal lengths (gap fills): substitutions only
            t = q.copy()
            for k in range(qlen):
                if rng.random() < float(rng.choice([0.0, 0.01, 0.05, 0.2])):
                    t[k] = (t[k] + rng.integers(1, 4)) % 4
            flag = int(rng.choice([0x08, 0x00]))
            eb = -1
        elif kind == 5:    # global, unequal lengths
            t = _mutate(rng, q, 0.03, 0.02)
            flag = int(rng.choice([0x08, 0x00]))
            eb = -1
        else:              # max_sw_mat skip (or just under it)
            t = _mutate(rng, q, 0.02, 0.01)
            flag = int(rng.choice([0x40, 0x00]))
            msm = max(1, len(t) * qlen - int(rng.integers(-2, 3)))
        out.append((np.array(p, np.int32), msm, q, np.ascontiguousarray(t, np.uint8), mat, w, eb, zd, flag))
    return out
