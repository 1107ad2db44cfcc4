"""GPU parity of the FM-index kernels (rank/extend, k-mer cache, SA locate, two-pass seeding).

Each test calls libmbg.so through its C ABI and compares with the compiled reference
(oracle/_ref/libmbref.so: mb_bwt_rank2a, mb_bwt_extend, mb_bwt_sa_batch, mb_seed_intv_batch) and
with the C restatement (oracle/liboracle.so).  Bit-exact integers throughout.
"""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx_chrm(gpu_ctx, chrm_prefix):
    gpu_ctx.idx_load(chrm_prefix, 10)
    return gpu_ctx


@pytest.fixture(scope="module")
def ref_chrm(chrm_prefix):
    from oracle import oracle as O
    return O.RefIndex(chrm_prefix + ".mbw")


@pytest.fixture(scope="module")
def orc_chrm(chrm_prefix):
    from oracle import oracle as O
    return O.OracleIndex(chrm_prefix + ".mbw")


def _edge_rows(n, primary):
    e = [0, 1, 2, n - 1, n, n + 1, primary - 1, primary, primary + 1, primary + 2]
    return sorted(set(v for v in e if 0 <= v <= n + 1))


def test_rank2a(ctx_chrm, ref_chrm):
    info = ctx_chrm.info()
    n, primary = info["seq_len"], info["primary"]
    rng = np.random.default_rng(7)
    k = rng.integers(0, n + 1, 20000).astype(np.uint64)
    l = np.minimum(k + rng.integers(0, 300, 20000).astype(np.uint64), n + 1)
    # l == k+1 at k == primary exercises the reference's shortcut (bwt.c:205-209)
    ed = _edge_rows(n, primary)
    ek = np.array([a for a in ed for b in ed if a <= b], np.uint64)
    el = np.array([b for a in ed for b in ed if a <= b], np.uint64)
    k = np.concatenate([k, ek, [primary]]).astype(np.uint64)
    l = np.concatenate([l, el, [primary + 1]]).astype(np.uint64)
    gk, gl = ctx_chrm.bwt_rank2a(k, l)
    for i in range(len(k)):
        rk, rl = ref_chrm.rank2a(k[i], l[i])
        assert np.array_equal(gk[i], rk) and np.array_equal(gl[i], rl), (i, k[i], l[i])


def test_extend(ctx_chrm, ref_chrm):
    rng = np.random.default_rng(8)
    info = ctx_chrm.info()
    n = info["seq_len"]
    ik = []
    for _ in range(5000):
        a = int(rng.integers(0, n + 1))
        s = int(rng.integers(0, min(200, n + 1 - a) + 1))
        b = int(rng.integers(0, n + 1 - s + 1))
        ik.append((a, b, s, 0))
    ik = np.array(ik, np.uint64)
    isb = rng.integers(0, 2, len(ik)).astype(np.uint8)
    ok = ctx_chrm.bwt_extend(ik, isb)
    for i in range(len(ik)):
        # is_back=1 ranks x[0]; is_back=0 ranks x[1]: give the reference the same interval
        r = ref_chrm.extend(ik[i], int(isb[i]))
        for c in range(4):
            assert tuple(int(v) for v in ok[i, c, :3]) == tuple(int(v) for v in r[c]), (i, c)


def test_kmer_cache(ctx_chrm, orc_chrm):
    g = ctx_chrm.pre_cache()
    import ctypes as C
    o = np.ctypeslib.as_array(C.cast(orc_chrm.b.contents.pre, C.POINTER(C.c_uint64)), shape=(4 << 20,)).reshape(-1, 4)
    assert np.array_equal(g[:, :3], o[:, :3])


def test_sa_all_rows(ctx_chrm, ref_chrm):
    n = ctx_chrm.info()["seq_len"]
    rows = np.arange(n + 1, dtype=np.uint64)
    g = ctx_chrm.bwt_sa_batch(rows)
    r = ref_chrm.sa(rows)
    assert np.array_equal(g, r)
    # the SA is a permutation of 0..n-1 plus the '$' row wrapped to 2^64-1+steps
    ok = g[g < n]
    assert len(np.unique(ok)) == len(ok)


def test_sa_random_order(ctx_chrm, ref_chrm):
    n = ctx_chrm.info()["seq_len"]
    rng = np.random.default_rng(9)
    rows = rng.integers(0, n + 1, 100000).astype(np.uint64)
    assert np.array_equal(ctx_chrm.bwt_sa_batch(rows), ref_chrm.sa(rows))


def _canon(a):
    a = np.asarray(a, np.uint64).reshape(-1, 4)
    if len(a) == 0:
        return a
    return a[np.lexsort((a[:, 3], a[:, 2], a[:, 1], a[:, 0]))]


def test_seed_chrm_reads(ctx_chrm, ref_chrm, chrm_reads):
    g = ctx_chrm.seed_intv_batch(chrm_reads, 19, 10)
    r = ref_chrm.seed_batch(chrm_reads, 19, 10)
    n_int = 0
    for i in range(len(chrm_reads)):
        assert np.array_equal(_canon(g[i]), _canon(r[i])), i
        n_int += len(g[i])
    assert n_int > len(chrm_reads)


@pytest.mark.parametrize("min_len,max_sub_occ", [(19, 10), (11, 3), (10, 10), (25, 50)])
def test_seed_params_and_edges(ctx_chrm, ref_chrm, chrm_reads, min_len, max_sub_occ):
    rng = np.random.default_rng(min_len)
    reads = [r.copy() for r in chrm_reads[:300]]
    for r in reads[::7]:                       # N bases (code 4) force the stage-1 skips
        r[rng.integers(0, len(r), 3)] = 4
    reads += [np.zeros(0, np.uint8), np.array([1, 2, 3], np.uint8), rng.integers(0, 4, 1000).astype(np.uint8)]
    g = ctx_chrm.seed_intv_batch(reads, min_len, max_sub_occ)
    r = ref_chrm.seed_batch(reads, min_len, max_sub_occ)
    for i in range(len(reads)):
        assert np.array_equal(_canon(g[i]), _canon(r[i])), i


def test_seed_cfg0(gpu_ctx):
    """Config-0 reference: 2,000 read pairs from the seeded generator."""
    from oracle import oracle as O
    from minibwa_amd import synth
    pre = os.path.join(GOLDEN, "cfg0")
    gpu_ctx.idx_load(pre, 10)
    ref = synth.reference(16569, seed=11, rep_frac=0.0)
    r1, r2 = synth.read_pairs(ref, 2000, seed=11)
    reads = list(r1) + list(r2)
    ri = O.RefIndex(pre + ".mbw")
    g = gpu_ctx.seed_intv_batch(reads)
    r = ri.seed_batch(reads)
    for i in range(len(reads)):
        assert np.array_equal(_canon(g[i]), _canon(r[i])), i


@pytest.mark.parametrize("pre_len", [0, 12, 14, 15])
def test_seed_cache_length_invariant(chrm_prefix, ref_chrm, chrm_reads, pre_len):
    """Seeding with no k-mer cache or a longer one than the reference's 10 gives the reference's
    intervals: a cached interval is the state stepwise extension reaches, and one below min_occ
    falls back to stepwise (bwt.c:403-406).  On chrM most 12- to 15-mers are absent, so the
    fallback is taken often; pass 2 (min_occ = size+1) exercises it with larger thresholds, and
    min_len 13 < 14, 15 runs the searches that start without a lookup."""
    from minibwa_amd import mbg
    reads = chrm_reads[:600]
    with mbg.Context(0) as c:
        c.idx_load(chrm_prefix, pre_len)
        for min_len, mso in ((19, 10), (13, 30)):
            g = c.seed_intv_batch(reads, min_len, mso)
            r = ref_chrm.seed_batch(reads, min_len, mso)
            for i in range(len(reads)):
                assert np.array_equal(_canon(g[i]), _canon(r[i])), (pre_len, min_len, i)


def test_seed_cfg0_cache14(gpu_ctx):
    """Config-0 reads against its index with a 14-mer cache."""
    from oracle import oracle as O
    from minibwa_amd import mbg, synth
    pre = os.path.join(GOLDEN, "cfg0")
    ref = synth.reference(16569, seed=11, rep_frac=0.0)
    r1, r2 = synth.read_pairs(ref, 1000, seed=13)
    reads = list(r1) + list(r2)
    ri = O.RefIndex(pre + ".mbw")
    with mbg.Context(0) as c:
        c.idx_load(pre, 14)
        g = c.seed_intv_batch(reads)
    r = ri.seed_batch(reads)
    for i in range(len(reads)):
        assert np.array_equal(_canon(g[i]), _canon(r[i])), i


def _long_reads(ref, rng, n):
    """Reads of 1-16 kb from chrM (either strand) at 0.1-15% substitutions, some with N runs and some
    with an unrelated stretch, where the segments' chains join the true chain late or not at all."""
    out = []
    for it in range(n):
        L = int(rng.integers(1024, min(16000, len(ref) - 1)))
        p = int(rng.integers(0, len(ref) - L))
        q = ref[p:p + L].copy()
        if it % 2:
            q = q[::-1].copy()
            q = np.where(q < 4, 3 - q, q).astype(np.uint8)
        sub = rng.random(L) < rng.choice([0.001, 0.005, 0.02, 0.08, 0.15])
        q[sub] = rng.integers(0, 4, int(sub.sum()))
        if it % 4 == 0:
            a = int(rng.integers(0, L - 50))
            q[a:a + int(rng.integers(1, 50))] = 4
        if it % 3 == 0:
            a = int(rng.integers(0, L - 700))
            q[a:a + 700] = rng.integers(0, 4, 700)
        out.append(q.astype(np.uint8))
    return out


@pytest.mark.parametrize("seg,side", [("512", "0"), ("512", "1"), ("100,3,8,2", "0"), ("64,1,2,1", "1"), ("0", "0")])
def test_seed_long_reads_segmented(chrm_prefix, ref_chrm, chrm_reads, seg, side):
    """Long reads: pass 1 splits each read of at least 2*len bases into segments searched in
    parallel, joined to the true chain by a fix-up, with a serial fallback for reads whose segment
    or fix-up tasks run out of room (MBG_SEED_SEG=len,segment slots,chain cap,fix-up slots; "0"
    is the unsegmented search).  Tiny capacities force the fallback on most reads.  MBG_SEED_SIDE=1
    runs the segment lanes on a second stream beside pass 1 proper.  Every read's intervals must
    equal the reference's mb_seed_intv_batch, mixed with short reads in one batch."""
    from tests.conftest import read_fasta, nt4
    from minibwa_amd import mbg
    _, seqs = read_fasta(os.path.join(GOLDEN, "chrM-human.fa.gz"))
    ref = nt4(seqs[0])
    reads = _long_reads(ref, np.random.default_rng(41), 120) + list(chrm_reads[:200])
    env = {"MBG_SEED_SEG": seg, "MBG_SEED_SIDE": side}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        with mbg.Context(0) as c:
            c.idx_load(chrm_prefix, 10)
            for min_len, mso in ((19, 10), (11, 4)):
                print("[seed-long] seg %s side %s min_len %d" % (seg, side, min_len), flush=True)
                g = c.seed_intv_batch(reads, min_len, mso)
                r = ref_chrm.seed_batch(reads, min_len, mso)
                for i in range(len(reads)):
                    assert np.array_equal(_canon(g[i]), _canon(r[i])), (seg, side, min_len, i, len(reads[i]))
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
