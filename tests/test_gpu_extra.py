"""GPU parity of mbg_update_extra_batch (mb_fix_cigar + mb_update_extra's scores, align.c:70-258)
against the oracle, which the CPU tests pin against the reference's own mb_update_extra: random
CIGARs with adjacent I/D runs, zero-length runs, leading gaps and low-complexity bases that make
indels slide, Ns, both strands, both gap models (q+e and the mb_log2 form), and hits from the
extension DP.
"""
import numpy as np
import pytest

from tests import brute

O = pytest.importorskip("oracle.oracle")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("log_gap", [0, 1])
def test_update_extra_parity(gpu_ctx, log_gap):
    rng = np.random.default_rng(200 + log_gap)
    mat = O.gen_nt4_mat(2, 8, 0, 1)
    hits, cigs, qs_, ts_ = [], [], [], []
    for it in range(4000):
        if it % 4 == 0:
            ql = int(rng.integers(20, 250))
            q = rng.integers(0, 4, ql).astype(np.uint8)
            t = q.copy()
            t[rng.random(ql) < 0.03] = rng.integers(0, 4)
            t = np.insert(t, int(rng.integers(0, ql)), rng.integers(0, 4, int(rng.integers(1, 5))))
            d = O.extz2(q, t, mat, 12, 2, 101, -1, 0, 0)
            cig = d["cigar"]
        else:
            cig, q, t = brute.random_alignment(rng, lowcomp=rng.random() < 0.5)
        ql = sum(c >> 4 for c in cig if c & 0xf in (0, 1))
        tl = sum(c >> 4 for c in cig if c & 0xf in (0, 2))
        hits.append(dict(qs=int(rng.integers(0, 50)), qe=0, ts=int(rng.integers(0, 500)), te=0, rev=int(rng.random() < 0.5)))
        hits[-1]["qe"] = hits[-1]["qs"] + ql
        hits[-1]["te"] = hits[-1]["ts"] + tl
        cigs.append(list(cig)), qs_.append(q[:ql]), ts_.append(t[:tl])
    jobs, fixed = gpu_ctx.update_extra(hits, cigs, qs_, ts_, mat, 12, 2, log_gap)
    n_changed = 0
    for i in range(len(hits)):
        want, wc = O.orc_update_extra(hits[i], cigs[i], qs_[i], ts_[i], mat, 12, 2, log_gap)
        got = {f: int(jobs[f][i]) for f in O.EXTRA_FIELDS}
        if not cigs[i]:
            continue
        assert got == want and fixed[i] == wc, (i, cigs[i], got, want, fixed[i], wc)
        n_changed += wc != cigs[i]
    assert n_changed > 500
