"""The served-mapping workload: read pools and a request schedule for mbg-serve-bench.

Requests arrive as a Poisson process at a given rate; a request is n read pairs of 2x150 bp with n
log-uniform in [1, max_pairs], or (a fraction long_frac of the requests) one single read of
long_len bases.  Pairs and long reads come from the synthetic reference with the workload's error
models (minibwa_amd/synth.py); the pools are FASTA (no qualities), one line per sequence, mates in
r1.fa and r2.fa under the same names, which the reference CLI reads as well.
"""
import os

import numpy as np

from minibwa_amd import synth

_ACGT = np.frombuffer(b"ACGTN", np.uint8)


def _write_fasta(path, seq, off, names):
    """seq: nt4 codes; record i is seq[off[i]:off[i+1]] under names[i]."""
    txt = _ACGT[np.minimum(np.asarray(seq, np.uint8), 4)]
    with open(path, "wb") as f:
        for i, nm in enumerate(names):
            f.write(b">" + nm.encode() + b"\n")
            f.write(txt[off[i]:off[i + 1]].tobytes())
            f.write(b"\n")


def write_pools(ref, work, n_pairs, n_long, seed=11, long_len=10000, prefix="s"):
    """r1.fa, r2.fa (n_pairs mates) and long.fa (n_long reads) under work; returns their paths."""
    r1, r2 = synth.read_pairs(ref, n_pairs, seed=seed, rlen=150, frag_mean=400, frag_sd=50, sub=0.01, indel=0.001)
    off = np.arange(n_pairs + 1, dtype=np.int64) * 150
    names = ["%sp_%d" % (prefix, i) for i in range(n_pairs)]
    p1, p2, pl = (os.path.join(work, x) for x in ("r1.fa", "r2.fa", "long.fa"))
    _write_fasta(p1, r1.reshape(-1), off, names)
    _write_fasta(p2, r2.reshape(-1), off, names)
    seq, loff, _ = synth.single_reads(ref, n_long, seed=seed + 2, lmin=long_len, lmax=long_len, sub=0.001, indel=0.0005)
    _write_fasta(pl, seq, loff, ["%sl_%d" % (prefix, i) for i in range(n_long)])
    return p1, p2, pl


def schedule(n_req, qps, seed=7, max_pairs=10000, long_frac=0.05):
    """[(arrival s, 'P' or 'L', n)]: Poisson arrivals at qps (all at 0 when qps <= 0)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / qps, n_req)) if qps > 0 else np.zeros(n_req)
    long_ = rng.random(n_req) < long_frac
    n = np.floor(np.exp(rng.uniform(0.0, np.log(max_pairs + 1), n_req))).astype(np.int64)
    n = np.clip(n, 1, max_pairs)
    return [(float(t[i]), "L" if long_[i] else "P", 1 if long_[i] else int(n[i])) for i in range(n_req)]


def write_schedule(path, sched):
    with open(path, "w") as f:
        for t, k, n in sched:
            f.write("%.6f %s %d\n" % (t, k, n))


def pool_prefix(path, n, out):
    """The first n records of a pool FASTA (two lines each) into out: the reads a run consumed."""
    with open(path) as f, open(out, "w") as g:
        for i, line in enumerate(f):
            if i >= 2 * n:
                break
            g.write(line)
