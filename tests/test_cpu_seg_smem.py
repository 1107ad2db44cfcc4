"""The segmented pass-1 SMEM search used for long reads, checked on the CPU with the oracle's serial
step (orc_smem, a restatement of mb_bwt_smem, bwt.c:280-305, pinned to the reference by the seeding
tests).  The serial search is a chain x0 = 0 -> x1 -> ... where each step from x emits at most one
SMEM and returns the next start.  The device splits a read of length L >= 2*S into K = L // S
segments [k*S, (k+1)*S) (the last one runs to L) and starts one lane per segment at k*S, which stops
at its first chain value >= (k+1)*S (its handoff) and records its chain values.  Segment 0 is exact.
A fix-up then walks the segments in order with the true chain value x (first: segment 0's handoff):
when x is one of segment k's chain values, segment k's lane has joined the true chain there (its
steps from x on are true steps, the ones before are not) and x jumps to its handoff; otherwise the
fix-up takes true steps itself until it lands on the chain or passes the segment.  This test runs
that procedure, with the same step function, and compares the SMEM set with the serial search's, on
long reads with errors, N runs and unrelated stretches (where chains do not join at once)."""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN, nt4, read_fasta

O = pytest.importorskip("oracle.oracle")
INF = 1 << 62


def serial(oi, q, min_len=19):
    out, x = [], 0
    while True:
        x, p = oi.smem_step(q, x, min_len)
        if p is not None:
            out.append(p)
        if x >= len(q):
            return out


def segmented(oi, q, S, min_len=19):
    L = len(q)
    K = L // S if L >= 2 * S else 1
    stop = [(k + 1) * S for k in range(K - 1)] + [L]
    chains, hand, recs = [], [], []
    for k in range(K):
        x, ch, rc = k * S, [], []
        while x < stop[k] and x < L:
            ch.append(x)
            x, p = oi.smem_step(q, x, min_len)
            if p is not None:
                rc.append(p)
        chains.append(ch), hand.append(x), recs.append(rc)
    conv = [0] + [INF] * (K - 1)
    fix, x, k, steps = [], hand[0], 1, 0
    while True:
        while k < K:
            if x >= stop[k]:
                conv[k] = INF
            elif x in chains[k]:
                conv[k], x = x, hand[k]
            else:
                break
            k += 1
        if k >= K or x >= L:
            break
        x, p = oi.smem_step(q, x, min_len)
        steps += 1
        if p is not None:
            fix.append(p)
    out = list(recs[0]) + fix
    for k in range(1, K):
        out += [p for p in recs[k] if (p[3] >> 32) >= conv[k]]
    return out, steps, sum(1 for c in conv[1:] if c == INF)


def test_segmented_pass1_equals_serial():
    oi = O.OracleIndex(os.path.join(GOLDEN, "chrM.mbw"), 10)
    _, seqs = read_fasta(os.path.join(GOLDEN, "chrM-human.fa.gz"))
    ref = nt4(seqs[0])
    rng = np.random.default_rng(5)
    n_fix = n_lost = 0
    for it in range(60):
        L = int(rng.integers(1200, 9000))
        p = int(rng.integers(0, len(ref) - L))
        q = ref[p:p + L].copy()
        if it % 2:
            q = q[::-1].copy()
            q = np.where(q < 4, 3 - q, q).astype(np.uint8)
        sub = rng.random(L) < rng.choice([0.001, 0.01, 0.05, 0.15])
        q[sub] = rng.integers(0, 4, int(sub.sum()))
        if it % 5 == 0:                      # an N run
            a = int(rng.integers(0, L - 40))
            q[a:a + int(rng.integers(1, 40))] = 4
        if it % 3 == 0:                      # an unrelated stretch: short chain steps
            a = int(rng.integers(0, L - 600))
            q[a:a + 600] = rng.integers(0, 4, 600)
        want = sorted(serial(oi, q))
        for S in (64, 256, 512):
            got, steps, lost = segmented(oi, q, S)
            assert sorted(got) == want, (it, S)
            n_fix += steps
            n_lost += lost
    assert n_fix > 0          # chains that did not join at once occurred and were repaired
