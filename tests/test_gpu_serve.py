"""Serving invariance: a request's results are the same whether its reads are processed alone or in a
mixed batch with other requests, and whether the batches run one after the other or concurrently
on two contexts (two streams) that share one resident index (mbg_idx_attach).

Per read the comparison covers the seed intervals (both passes), the located positions of its
anchor rows, and every extension job with its DP result and CIGAR -- all integers, compared exactly.
"""
import os
import threading

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

RES_FIELDS = ("max", "zdropped", "max_q", "max_t", "mqe", "mqe_t", "mte", "mte_q", "score", "reach_end", "n_cigar", "cells")


def _run(ctx, reads, opt):
    """One step over `reads` (list of uint8 arrays) from device memory; returns per-read results."""
    from minibwa_amd import mbg
    n = len(reads)
    lens = np.array([len(r) for r in reads], np.int32)
    off = np.concatenate([[0], np.cumsum(lens.astype(np.int64))])
    seq = np.concatenate(reads).astype(np.uint8)
    d_seq, d_off, d_len = ctx.to_device(seq), ctx.to_device(off), ctx.to_device(lens)
    st, out = ctx.map_step_dev(n, d_off.ptr, d_len.ptr, d_seq.ptr, int(lens.max()), opt)
    r = ctx.step_results(st, out)
    per = [{"seeds": set(), "pos": [], "jobs": {}} for _ in range(n)]
    for rec in r["recs"]:
        per[int(rec["rid"])]["seeds"].add((int(rec["x0"]), int(rec["x1"]), int(rec["size"]), int(rec["info"]), int(rec["pass"])))
    for p, k in zip(r["pos"], r["pos_rec"]):
        per[int(r["recs"][int(k)]["rid"])]["pos"].append(int(p))
    for j, m, res in zip(r["jobs"], r["meta"], r["res"]):
        cg = r["cigar"][int(j["cigar_off"]):int(j["cigar_off"]) + int(res["n_cigar"])]
        per[int(m["rid"])]["jobs"][int(m["side"])] = (
            int(j["qlen"]), int(j["tlen"]), int(j["w"]), int(j["flag"]), int(m["rev"]), int(m["tid"]), int(m["tpos"]),
            tuple(int(res[f]) for f in RES_FIELDS), tuple(int(c) for c in cg))
    for p in per:
        p["pos"].sort()
    return per


@pytest.fixture(scope="module")
def cfg0_two_contexts():
    from minibwa_amd import mbg
    a = mbg.Context(0)
    a.idx_load(os.path.join(GOLDEN, "cfg0"), 10)
    b = mbg.Context(0)
    b.idx_attach(a)
    yield a, b
    b.close()
    a.close()


def _requests():
    from minibwa_amd import synth
    ref = synth.reference(16569, seed=11, rep_frac=0.0)
    r1, r2 = synth.read_pairs(ref, 400, seed=17)
    reads = []
    for i in range(400):
        reads += [r1[i], r2[i]]
    return reads[:100], reads[100:]            # request A (50 pairs) and the rest of the traffic


def test_alone_equals_mixed(cfg0_two_contexts):
    from minibwa_amd import mbg
    a, b = cfg0_two_contexts
    opt = mbg.adaptive_opt(150)
    A, other = _requests()
    alone = _run(b, A, opt)                      # the attached context, alone
    mix = _run(a, other[:200] + A + other[200:], opt)
    assert sum(len(p["jobs"]) for p in alone) > 0
    for i in range(len(A)):
        assert alone[i] == mix[200 + i], i


def test_concurrent_contexts_equal_sequential(cfg0_two_contexts):
    from minibwa_amd import mbg
    a, b = cfg0_two_contexts
    opt = mbg.adaptive_opt(150)
    A, other = _requests()
    seq_a, seq_o = _run(b, A, opt), _run(a, other, opt)
    got = {}

    def work(key, ctx, reads):
        got[key] = _run(ctx, reads, opt)

    for _ in range(3):
        t1 = threading.Thread(target=work, args=("A", b, A))
        t2 = threading.Thread(target=work, args=("O", a, other))
        t1.start(), t2.start()
        t1.join(), t2.join()
        assert got["A"] == seq_a and got["O"] == seq_o
