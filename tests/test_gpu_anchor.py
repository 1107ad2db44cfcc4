"""GPU parity of mbg_anchor_batch against the oracle's mb_anchor (seed.c:268-339), which the CPU tests pin
against the reference's own function: anchors, their order (klib's radix sort leaves equal positions in
an order of its own) and the dedup, on a repeat-rich 4-contig index (lists of more than 64 intervals and
of more than 64 anchors, intervals of more than max_occ rows) and on chrM with its golden reads.  The
intervals come back sorted and deduplicated; their order is the closed form of mb_seed_sort_dedup:
by x[0], then size descending, then info.
"""
import numpy as np
import pytest

from minibwa_amd import mbg

O = pytest.importorskip("oracle.oracle")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rep(tmp_path_factory):
    from minibwa_amd import index, synth
    ref = synth.repeat_reference()
    pre = str(tmp_path_factory.mktemp("anc") / "rep")
    parts = index.build(ref, pre, n_ctg=4, ctg_prefix="rc", device="cpu")
    ctx = mbg.Context(0)
    ctx.idx_load(pre, 10)
    oi = O.OracleIndex(pre + ".mbw", pre_len=10)
    yield ctx, oi, ref, parts["ctg_len"]
    oi.close()
    ctx.close()


def _reads(ref):
    from minibwa_amd import synth
    r1, r2 = synth.read_pairs(ref, 400, seed=9)
    reads = [r for r in r1] + [r for r in r2]
    lr, lo, ll = synth.single_reads(ref, 24, seed=4, lmin=2000, lmax=8000, sub=0.04)
    return reads + [lr[lo[i]:lo[i] + ll[i]] for i in range(len(ll))]


def _sort_dedup_closed_form(u):
    u = np.asarray(u, np.uint64).reshape(-1, 4)
    if len(u) == 0:
        return u
    o = np.lexsort((u[:, 3], np.iinfo(np.uint64).max - u[:, 2], u[:, 0]))
    u = u[o]
    keep = np.ones(len(u), bool)
    keep[1:] = (u[1:, 0] != u[:-1, 0]) | (u[1:, 2] != u[:-1, 2]) | (u[1:, 3] != u[:-1, 3])
    return u[keep]


def _check(ctx, oi, ctg_len, reads, ivs, max_occ):
    anc, us = ctx.anchor(ivs, [len(r) for r in reads], max_occ)
    n_big = 0
    for i, r in enumerate(reads):
        want = O.orc_anchor(oi, ctg_len, ivs[i], len(r), max_occ)
        got = anc[i]
        assert len(got) == len(want) and (got == want).all(), (i, max_occ, len(got), len(want))
        assert (us[i] == _sort_dedup_closed_form(ivs[i])).all(), (i, max_occ)
        n_big += len(got) > 64
    return n_big


@pytest.mark.parametrize("max_occ", [250, 20, 3, 1])
def test_anchor_repeat_index(rep, max_occ):
    ctx, oi, ref, ctg_len = rep
    reads = _reads(ref)
    ivs = ctx.seed_intv_batch(reads, 19, 10)
    assert sum(len(u) > 64 for u in ivs) > 5
    assert sum(int((np.asarray(u)[:, 2] > max_occ).sum()) for u in ivs if len(u)) > 50
    n_big = _check(ctx, oi, ctg_len, reads, ivs, max_occ)
    if max_occ >= 20:
        assert n_big > 20


def test_anchor_chrm(gpu_ctx, chrm_prefix, chrm_reads):
    gpu_ctx.idx_load(chrm_prefix, 10)
    oi = O.OracleIndex(chrm_prefix + ".mbw", pre_len=10)
    reads = chrm_reads[:2000]
    ivs = gpu_ctx.seed_intv_batch(reads, 19, 10)
    ctg_len = np.array([gpu_ctx.info()["tot_len"]], np.uint64)
    _check(gpu_ctx, oi, ctg_len, reads, ivs, 250)
    oi.close()


def test_anchor_edges(rep):
    """Reads without intervals, an empty batch, max_occ 0, shuffled and duplicated interval lists."""
    ctx, oi, ref, ctg_len = rep
    reads = _reads(ref)[:60]
    ivs = ctx.seed_intv_batch(reads, 19, 10)
    rng = np.random.default_rng(1)
    mixed = []
    for u in ivs:
        u = np.asarray(u, np.uint64).reshape(-1, 4)
        if len(u):
            u = np.concatenate([u, u[rng.integers(0, len(u), len(u) // 2)]])[rng.permutation(len(u) + len(u) // 2)]
        mixed.append(u)
    mixed[3] = np.zeros((0, 4), np.uint64)
    _check(ctx, oi, ctg_len, reads, mixed, 250)
    anc, us = ctx.anchor(mixed, [len(r) for r in reads], 0)
    assert all(len(a) == 0 for a in anc)
    anc, us = ctx.anchor([], [], 250)
    assert anc == [] and us == []


def test_anchor_errors(rep):
    ctx, oi, ref, ctg_len = rep
    seq_len = ctx.info()["seq_len"]
    bad = np.array([[seq_len - 2, 0, 5, (3 << 32) | 40]], np.uint64)        # rows past the index
    with pytest.raises(mbg.MbgError, match="invalid"):
        ctx.anchor([bad], [100], 250)
    with pytest.raises(mbg.MbgError, match="invalid"):
        ctx.anchor([np.zeros((0, 4), np.uint64)], [100], -1)


def test_anchor_dev_short_slots_twice(rep):
    """mbg_anchor_batch_dev twice on one workspace, one read given a slot fewer than it needs: that read
    gets n_anc -1 on both calls (its rows are all rewritten, so the second call's SA walk never sees a
    value the first call located there), every other read the host entry point's anchors."""
    import ctypes as C
    ctx, oi, ref, ctg_len = rep
    P = C.c_void_p
    reads = _reads(ref)[:300]
    ivs = ctx.seed_intv_batch(reads, 19, 10)
    n = len(reads)
    want, _ = ctx.anchor(ivs, [len(r) for r in reads], 250)
    us = [np.asarray(u, np.uint64).reshape(-1, 4) for u in ivs]
    uoff = np.zeros(n + 1, np.int64)
    np.cumsum([len(u) for u in us], out=uoff[1:])
    flat = np.ascontiguousarray(np.concatenate(us))
    slots = mbg.anchor_slots(us, 250)

    def need(u):                                   # rows located once per group of equal (x[0], size)
        g = np.unique(u[:, [0, 2]], axis=0) if len(u) else np.zeros((0, 2), np.uint64)
        return int(np.minimum(g[:, 1], 250).sum())
    short = next(i for i in range(n) if need(us[i]) >= 3)
    slots[short] = need(us[short]) - 1
    aoff = np.zeros(n + 1, np.int64)
    np.cumsum(slots, out=aoff[1:])
    ns = int(aoff[-1])
    d_u0, d_u = ctx.to_device(flat), ctx.to_device(flat)
    d_uoff, d_ql, d_aoff = ctx.to_device(uoff), ctx.to_device(np.array([len(r) for r in reads], np.int32)), ctx.to_device(aoff)
    d_rows, d_anc = ctx.to_device(np.zeros(ns, np.uint64)), ctx.to_device(np.zeros(ns, mbg.ANCHOR_DTYPE))
    d_nanc, d_nu = ctx.to_device(np.zeros(n, np.int64)), ctx.to_device(np.zeros(n, np.int64))
    for call in range(2):
        mbg.check(ctx.L.mbg_stream_copy(ctx.h, None, P(d_u.ptr), P(d_u0.ptr), flat.nbytes), "mbg_stream_copy")
        mbg.check(ctx.L.mbg_anchor_batch_dev(ctx.h, None, n, ns, P(d_u.ptr), P(d_uoff.ptr), P(d_ql.ptr), 250, P(d_rows.ptr),
                                             P(d_anc.ptr), P(d_aoff.ptr), P(d_nanc.ptr), P(d_nu.ptr)), "mbg_anchor_batch_dev")
        nanc, anc = d_nanc.numpy(), d_anc.numpy()
        assert nanc[short] == -1, call
        for i in range(n):
            if i != short:
                assert nanc[i] == len(want[i]) and (anc[aoff[i]:aoff[i] + nanc[i]] == want[i]).all(), (call, i)
    for d in (d_u0, d_u, d_uoff, d_ql, d_aoff, d_rows, d_anc, d_nanc, d_nu):
        d.free()
