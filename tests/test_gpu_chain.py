"""GPU parity of mbg_lchain_batch against the oracle's mb_lchain_dp (lchain.c:145-240), which the CPU
tests pin against the reference's own: chain scores and anchor counts, the order of chains (klib's radix
sort of equal scores and equal start positions) and the chained anchors, on the device's own anchors of
short and long reads on the repeat-rich index, over the short-read, long-read and re-chaining option
sets and tight skip/iteration limits.
"""
import numpy as np
import pytest

from minibwa_amd import mbg
from tests.test_gpu_anchor import _reads, rep  # noqa: F401  (module fixture)

O = pytest.importorskip("oracle.oracle")

pytestmark = pytest.mark.gpu


def _opt(p):
    return mbg.ChainOpt(max_dist_x=p["max_dist_x"], max_dist_y=p["max_dist_y"], bw=p["bw"], max_skip=p["max_skip"],
                        max_iter=p["max_iter"], min_sc=p["min_sc"], pen_gap=p["pen_gap"])


@pytest.fixture(scope="module")
def anchors(rep):
    ctx, oi, ref, ctg_len = rep
    reads = _reads(ref)
    ivs = ctx.seed_intv_batch(reads, 19, 10)
    anc, _ = ctx.anchor(ivs, [len(r) for r in reads], 250)
    return anc


@pytest.mark.parametrize("pi", range(len(O.CHAIN_PARAM_SETS)))
def test_lchain_parity(rep, anchors, pi):
    ctx, oi, ref, ctg_len = rep
    p = O.CHAIN_PARAM_SETS[pi]
    got = ctx.lchain(anchors, _opt(p))
    n_multi = 0
    for i, a in enumerate(anchors):
        u, b = O.orc_lchain(ctg_len, a, p)
        gu, gb = got[i]
        assert len(gu) == len(u) and (gu == u).all(), (i, pi, len(gu), len(u))
        assert len(gb) == len(b) and (gb == b).all(), (i, pi)
        n_multi += len(u) > 1
    assert n_multi > 50
    assert max(len(a) for a in anchors) > 1000


def test_lchain_edges(rep, anchors):
    """Reads without anchors, one anchor, chains below min_sc, an empty batch, a bad contig."""
    ctx, oi, ref, ctg_len = rep
    p = O.CHAIN_SR
    sel = [anchors[0][:0], anchors[1][:1], anchors[2], anchors[3][:0], anchors[-1]]
    got = ctx.lchain(sel, _opt(p))
    for i, a in enumerate(sel):
        u, b = O.orc_lchain(ctg_len, a, p)
        assert (got[i][0] == u).all() and len(got[i][0]) == len(u) and (got[i][1] == b).all(), i
    high = dict(p, min_sc=1 << 20)
    assert all(len(u) == 0 and len(b) == 0 for u, b in ctx.lchain(sel, _opt(high)))
    assert ctx.lchain([], _opt(p)) == []
    bad = anchors[2].copy()
    bad["sid"][0] = 2 * 9                                    # contig 9 of a 4-contig reference
    with pytest.raises(mbg.MbgError, match="invalid"):
        ctx.lchain([bad], _opt(p))
