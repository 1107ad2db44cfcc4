"""Seeded synthetic workloads (bindings to minibwa_amd/libmbsim.so, source sim/mbsim.c).

Every generator is splitmix64-seeded, so the same seed and sizes give the same bytes here and on
the GPU host.  Sequences are nt4 uint8 arrays (A,C,G,T = 0..3).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBSIM = os.path.join(HERE, "libmbsim.so")
_L = None


def _lib():
    global _L
    if _L is None:
        if not os.path.exists(LIBSIM):
            raise RuntimeError(f"{LIBSIM} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIBSIM)
        P = C.c_void_p
        L.mbs_ref.argtypes = [C.c_uint64, C.c_int64, C.c_double, C.c_int32, C.c_int32, C.c_double, P]
        L.mbs_pairs.argtypes = [C.c_uint64, P, C.c_int64, C.c_int64, C.c_int32, C.c_double, C.c_double, C.c_double,
                                C.c_double, P, P, P]
        L.mbs_single.restype = C.c_int64
        L.mbs_single.argtypes = [C.c_uint64, P, C.c_int64, C.c_int64, C.c_int32, C.c_int32, C.c_double, C.c_double, P,
                                 C.c_int64, P]
        L.mbs_dp_jobs.restype = C.c_int64
        L.mbs_dp_jobs.argtypes = [C.c_uint64, C.c_int64, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_int32,
                                  C.c_int, P, C.c_int64, P, P, C.c_int64, P]
        L.mbs_write_fastq.argtypes = [C.c_char_p, C.c_char_p, C.c_int64, P, P, P, C.c_int]
        L.mbs_write_fasta.argtypes = [C.c_char_p, C.c_char_p, P, C.c_int64, C.c_int32]
        _L = L
    return _L


def _p(a):
    return C.c_void_p(a.ctypes.data)


def reference(length, seed=11, rep_frac=0.05, rep_min=300, rep_max=6000, rep_div=0.02) -> np.ndarray:
    """Uniform ACGT of `length` bases with `rep_frac` of it overwritten by diverged segment copies."""
    out = np.empty(length, np.uint8)
    _lib().mbs_ref(seed, length, rep_frac, rep_min, rep_max, rep_div, _p(out))
    return out


def repeat_reference(seed=5, n_copies=320, spacing=900, elem_len=400, elem_div=0.001):
    """A repeat-rich reference: 300 kb of synth.reference (10% diverged segment copies) cut into
    n_copies stretches of `spacing` bases, each followed by a copy of one random elem_len-base element
    with elem_div substitutions -- reads on it get intervals of several hundred rows."""
    rng = np.random.default_rng(seed)
    base = reference(max(300_000, n_copies * spacing), seed=seed, rep_frac=0.1, rep_min=300, rep_max=3000, rep_div=0.01)
    elem = rng.integers(0, 4, elem_len).astype(np.uint8)
    parts = []
    for c in range(n_copies):
        cp = elem.copy()
        cp[rng.random(elem_len) < elem_div] = rng.integers(0, 4)
        parts += [base[c * spacing:(c + 1) * spacing], cp]
    return np.concatenate(parts).astype(np.uint8)


def read_pairs(ref, n_pairs, seed=11, rlen=150, frag_mean=400.0, frag_sd=50.0, sub=0.01, indel=0.001, truth=False):
    """(r1, r2) as (n_pairs, rlen) uint8 arrays; with truth=True also (n_pairs, 2) int64."""
    ref = np.ascontiguousarray(ref, np.uint8)
    r1 = np.empty((n_pairs, rlen), np.uint8)
    r2 = np.empty((n_pairs, rlen), np.uint8)
    tr = np.empty((n_pairs, 2), np.int64)
    rc = _lib().mbs_pairs(seed, _p(ref), len(ref), n_pairs, rlen, frag_mean, frag_sd, sub, indel, _p(r1), _p(r2), _p(tr))
    if rc != 0:
        raise RuntimeError("mbs_pairs failed")
    return (r1, r2, tr) if truth else (r1, r2)


def single_reads(ref, n, seed=11, lmin=100, lmax=20000, sub=0.001, indel=0.0005):
    """(bases, offsets[n+1], lengths[n]) of n single-end reads with log-uniform lengths."""
    ref = np.ascontiguousarray(ref, np.uint8)
    cap = int(n * (lmax + 1))
    cap = min(cap, int(n * np.exp((np.log(lmin) + np.log(lmax)) / 2) * 4) + lmax * 4)
    while True:
        out = np.empty(cap, np.uint8)
        ln = np.empty(n, np.int32)
        tot = _lib().mbs_single(seed, _p(ref), len(ref), n, lmin, lmax, sub, indel, _p(out), cap, _p(ln))
        if tot >= 0:
            break
        cap *= 2
    off = np.zeros(n + 1, np.int64)
    np.cumsum(ln, out=off[1:])
    return out[:tot], off, ln


def dp_jobs(n, seed=11, qmin=100, qmax=300, sub=0.02, indel=0.005, flank=50, two_sided=False):
    """Config-2 style DP jobs -> (qbuf, qoff, qlen, tbuf, toff, tlen)."""
    qcap = n * qmax + 16
    tcap = n * (qmax * 3 + 2 * flank) + 16
    qb = np.empty(qcap, np.uint8)
    tb = np.empty(tcap, np.uint8)
    ql = np.empty(n, np.int32)
    tl = np.empty(n, np.int32)
    tot = _lib().mbs_dp_jobs(seed, n, qmin, qmax, sub, indel, flank, int(two_sided), _p(qb), qcap, _p(ql), _p(tb), tcap,
                             _p(tl))
    if tot < 0:
        raise RuntimeError("mbs_dp_jobs failed")
    qoff = np.zeros(n + 1, np.int64)
    np.cumsum(ql, out=qoff[1:])
    toff = np.zeros(n + 1, np.int64)
    np.cumsum(tl, out=toff[1:])
    return qb[:qoff[-1]], qoff, ql, tb[:toff[-1]], toff, tl


def write_fasta(path, ref, n_ctg=1, prefix="chr"):
    ref = np.ascontiguousarray(ref, np.uint8)
    if _lib().mbs_write_fasta(path.encode(), prefix.encode(), _p(ref), len(ref), n_ctg) != 0:
        raise IOError(path)


def write_fastq(path, seq, off, lens, prefix="r", mate=0):
    seq = np.ascontiguousarray(seq, np.uint8)
    off = np.ascontiguousarray(off, np.int64)
    lens = np.ascontiguousarray(lens, np.int32)
    if _lib().mbs_write_fastq(path.encode(), prefix.encode(), len(lens), _p(seq), _p(off), _p(lens), mate) != 0:
        raise IOError(path)


def write_pairs_fastq(path1, path2, r1, r2, prefix="p"):
    n, rl = r1.shape
    off = np.arange(n + 1, dtype=np.int64) * rl
    lens = np.full(n, rl, np.int32)
    write_fastq(path1, r1.reshape(-1), off, lens, prefix, 1)
    write_fastq(path2, r2.reshape(-1), off, lens, prefix, 2)


def ref_to_l2b_arrays(ref: np.ndarray, n_ctg=1):
    """(ctg_len, pac) of the .l2b layout for an N-free nt4 reference split into n_ctg equal contigs
    (the split write_fasta uses)."""
    L = len(ref)
    per = (L + n_ctg - 1) // n_ctg
    ctg_len = np.array([min(per, L - c * per) for c in range(n_ctg)], np.uint64)
    n_pac = (L + 31) // 32
    padded = np.zeros(n_pac * 32, np.uint64)
    padded[:L] = ref
    sh = (np.arange(32, dtype=np.uint64) * 2)
    pac = (padded.reshape(n_pac, 32) << sh).sum(axis=1, dtype=np.uint64)
    return ctg_len, pac


def ll_jobs(n, seed=11, read_len=150, window=550, sub=0.01, frac_in=0.8):
    """Mate-rescue shaped ksw_ll jobs: a read of read_len bases against a window of `window` bases
    (insert-size range + read length), holding a copy of the read with `sub` substitutions in a
    fraction frac_in of the jobs -> (qbuf, qoff, qlen, tbuf, toff, tlen)."""
    rng = np.random.default_rng(seed)
    q = rng.integers(0, 4, (n, read_len), dtype=np.uint8)
    t = rng.integers(0, 4, (n, window), dtype=np.uint8)
    has = np.nonzero(rng.random(n) < frac_in)[0]
    st = rng.integers(0, window - read_len + 1, len(has))
    cp = q[has].copy()
    mut = rng.random(cp.shape) < sub
    cp[mut] = (cp[mut] + rng.integers(1, 4, int(mut.sum()), dtype=np.uint8)) % 4
    t[has[:, None], st[:, None] + np.arange(read_len)] = cp
    qoff = np.arange(n + 1, dtype=np.int64) * read_len
    toff = np.arange(n + 1, dtype=np.int64) * window
    return (q.reshape(-1), qoff, np.full(n, read_len, np.int32), t.reshape(-1), toff, np.full(n, window, np.int32))


def rescue_cands(ref, n, seed=11, read_len=150, window=550, sub=0.01, frac_in=0.8):
    """Mate-rescue candidates on one contig: a read_len-base mate (a copy of the reference with `sub`
    substitutions in a fraction frac_in of them, unrelated sequence otherwise) and a window of
    `window` bases around where the insert-size model puts it -> (reads [n, read_len], ts, te)."""
    rng = np.random.default_rng(seed)
    ref = np.asarray(ref, np.uint8)
    p = rng.integers(window, len(ref) - 2 * window, n)
    reads = ref[p[:, None] + np.arange(read_len)]
    mut = rng.random(reads.shape) < sub
    reads[mut] = (reads[mut] + rng.integers(1, 4, int(mut.sum()), dtype=np.uint8)) % 4
    other = rng.random(n) >= frac_in
    reads[other] = rng.integers(0, 4, (int(other.sum()), read_len), dtype=np.uint8)
    ts = p - rng.integers(0, window - read_len + 1, n)
    return reads, ts.astype(np.int64), (ts + window).astype(np.int64)
