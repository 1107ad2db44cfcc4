"""GPU parity of mbg_matesw_batch against the oracle's mate-rescue candidate: mb_ungap's vote, the
window narrowing and mb_matesw_align with its mb_fix_cigar + mb_update_extra (pe.c:209-312,342-349),
which the CPU tests pin against the reference's own functions.  Windows come from the chrM index resident
in HBM (l2b_getseq on the device); reads are mutated chrM segments (substitutions, indels, Ns) and
unrelated sequence, over the methylation modes and several option sets.
"""
import os

import numpy as np
import pytest

from minibwa_amd import mbg
from tests.conftest import GOLDEN, nt4, read_fasta

O = pytest.importorskip("oracle.oracle")

pytestmark = pytest.mark.gpu

EZ = ("max", "zdropped", "max_q", "max_t", "mqe", "mqe_t", "mte", "mte_q", "score", "reach_end", "n_cigar")
OPTS = [dict(a=2, b=8, b_ts=0, b_ambi=1, q=12, e=2, bw=100, zdrop=100, end_bonus=5, min_dp_max=30, min_len=19),
        dict(a=1, b=4, b_ts=0, b_ambi=1, q=6, e=1, bw=50, zdrop=60, end_bonus=0, min_dp_max=20, min_len=19),
        dict(a=2, b=8, b_ts=4, b_ambi=1, q=12, e=2, bw=30, zdrop=100, end_bonus=5, min_dp_max=40, min_len=19),
        dict(a=2, b=6, b_ts=0, b_ambi=1, q=8, e=3, bw=100, zdrop=400, end_bonus=10, min_dp_max=30, min_len=25)]


@pytest.fixture(scope="module")
def chrm(gpu_ctx, chrm_prefix):
    gpu_ctx.idx_load(chrm_prefix, 10)
    _, seqs = read_fasta(os.path.join(GOLDEN, "chrM-human.fa.gz"))
    return gpu_ctx, nt4(seqs[0])


def _mutate(rng, s):
    t = list(s)
    for _ in range(int(rng.integers(0, 4))):
        p = int(rng.integers(0, max(1, len(t))))
        if rng.random() < 0.5:
            del t[p:p + int(rng.integers(1, 6))]
        else:
            t[p:p] = list(rng.integers(0, 4, int(rng.integers(1, 6))))
    t = np.array(t, np.uint8)
    sub = rng.random(len(t)) < rng.choice([0.0, 0.01, 0.05, 0.15])
    t[sub] = rng.integers(0, 4, int(sub.sum()))
    if rng.random() < 0.1:
        t[rng.random(len(t)) < 0.02] = 4
    return t


def _candidates(rng, ref, n):
    L = len(ref)
    reads, ts, te = [], [], []
    for it in range(n):
        ql = int(rng.integers(20, 300))
        p = int(rng.integers(0, L - ql))
        r = _mutate(rng, ref[p:p + ql]) if rng.random() < 0.85 else rng.integers(0, 4, ql).astype(np.uint8)
        if len(r) < 8:
            r = rng.integers(0, 4, 30).astype(np.uint8)
        s = p - int(rng.integers(0, 600))
        e = s + len(r) + int(rng.integers(1, 800))
        if it % 50 == 0:
            s, e = -40, len(r) + 100              # clipped at the contig start
        if it % 50 == 1:
            s, e = L - len(r) - 300, L + 200      # clipped at the end
        reads.append(r), ts.append(s), te.append(e)
    return reads, np.array(ts, np.int64), np.array(te, np.int64)


def _compare(res, cigars, reads, ts, te, ref, p, mt, min_sc):
    mat_dp = O.gen_nt4_mat(int(p[0]), int(p[1]), int(p[2]), int(p[3]), mt)
    n_res = 0
    for i, r in enumerate(reads):
        s, e = max(0, int(ts[i])), min(len(ref), int(te[i]))
        got = res[i]
        if e - s <= len(r):
            assert got["aligned"] == 0 and got["rescued"] == 0, i
            continue
        d = O.orc_matesw_cand(p, r, ref[s:e], int(min_sc[i]), mt)
        ctx = (i, len(r), s, e)
        assert (int(got["max_ug"]), int(got["max_i"]), int(got["n_good"]), int(got["n_kmer"])) == d["ug"], ctx
        assert (int(got["ts2"]), int(got["te2"])) == d["win2"] and int(got["aligned"]) == d["aligned"], ctx
        assert {f: int(got["ll"][f]) for f in O.LL_FIELDS} == d["ll"], ctx
        assert int(got["ran_dp"]) == d["ran_dp"], ctx
        assert {f: int(got["ez"][f]) for f in EZ} == d["ez"], ctx
        assert int(got["rescued"]) == d["rescued"], ctx
        assert int(got["dp_max2"]) == d["dp_max2"], ctx
        if not d["rescued"]:
            assert cigars[i] == d["cigar"], ctx
            continue
        # the rescued hit after mb_fix_cigar + mb_update_extra (pe.c:309), from the oracle
        ts2 = d["win2"][0]
        hit = dict(qs=d["qs"], qe=d["qe"], ts=d["ts"], te=d["te"], rev=0)
        x, xc = O.orc_update_extra(hit, d["cigar"], r[d["qs"]:], ref[s + ts2 + d["ts"]:e], mat_dp, int(p[4]), int(p[5]), 0)
        assert cigars[i] == xc, ctx
        assert {f: int(got[f]) for f in ("qs", "qe", "ts", "te", "n_cigar", "blen", "mlen", "n_ambi", "dp_max")} == \
            {f: x[f] for f in ("qs", "qe", "ts", "te", "n_cigar", "blen", "mlen", "n_ambi", "dp_max")}, ctx
        n_res += 1
    return n_res


@pytest.mark.parametrize("mt", [0, 1, 2])
@pytest.mark.parametrize("opt_i", range(len(OPTS)))
def test_matesw_parity(chrm, mt, opt_i):
    ctx, ref = chrm
    rng = np.random.default_rng(10 * mt + opt_i)
    reads, ts, te = _candidates(rng, ref, 300)
    min_sc = rng.choice([0, 20, 40], len(reads))
    o = OPTS[opt_i]
    opt = mbg.MateswOpt(**o, mt=mt)
    res, cigars = ctx.matesw(reads, np.zeros(len(reads), np.int32), ts, te, min_sc, opt)
    p = O.rescue_params(**o)
    n_res = _compare(res, cigars, reads, ts, te, ref, p, mt, min_sc)
    assert n_res > 60
    assert (res["ts2"] > 0).sum() > 20                 # narrowed windows occur


def test_matesw_large_batch(chrm):
    """A serving-size batch of 150 bp mates against 550 bp windows (default options)."""
    ctx, ref = chrm
    rng = np.random.default_rng(99)
    n = 20000
    reads, ts, te = [], [], []
    for it in range(n):
        p = int(rng.integers(0, len(ref) - 150))
        r = _mutate(rng, ref[p:p + 150]) if it % 4 else rng.integers(0, 4, 150).astype(np.uint8)
        s = max(0, p - int(rng.integers(0, 400)))
        reads.append(r), ts.append(s), te.append(s + 550)
    ts, te = np.array(ts, np.int64), np.array(te, np.int64)
    min_sc = np.full(n, 20)
    opt = mbg.MateswOpt(**OPTS[0], mt=0)
    res, cigars = ctx.matesw(reads, np.zeros(n, np.int32), ts, te, min_sc, opt)
    idx = rng.choice(n, 800, replace=False)
    p = O.rescue_params(**OPTS[0])
    _compare(res[idx], [cigars[i] for i in idx], [reads[i] for i in idx], ts[idx], te[idx], ref, p, 0, min_sc[idx])
    assert res["rescued"].mean() > 0.5


def test_matesw_errors(chrm):
    ctx, ref = chrm
    opt = mbg.MateswOpt(**OPTS[0], mt=0)
    r = [ref[100:250]]
    with pytest.raises(mbg.MbgError, match="invalid"):
        ctx.matesw(r, np.array([5]), np.array([0]), np.array([600]), np.array([0]), opt)     # no contig 5
    with pytest.raises(mbg.MbgError, match="outside"):
        ctx.matesw(r, np.array([0]), np.array([0]), np.array([20000]), np.array([0]), opt)   # window > 16384
    o = dict(OPTS[0], q=0)
    with pytest.raises(mbg.MbgError, match="outside"):
        ctx.matesw(r, np.array([0]), np.array([0]), np.array([600]), np.array([0]), mbg.MateswOpt(**o, mt=0))
