"""Pins the oracle's restatement of mb_align_pair (align.c:329-390) against the reference's own static
function, exported by oracle/ref_align_shim.c (align.c compiled where it lies), on seeded calls that
reach every branch (tests/align_cases.py); and the fast paths against their definitions."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.align_cases import cases

pytestmark = pytest.mark.skipif(not O.have_ref_pe(), reason="oracle/_ref not built (no reference sources)")


def test_align_pair_vs_reference():
    branches = {"fast": 0, "dp": 0, "skip": 0}
    for i, (p, msm, q, t, mat, w, eb, zd, flag) in enumerate(cases(5, 3000)):
        got = O.orc_align_pair(p, msm, q, t, mat, w, eb, zd, flag)
        want = O.ref_align_pair(p, msm, q, t, mat, w, eb, zd, flag)
        cells = got.pop("cells")
        assert cells >= 0
        assert got == want, f"case {i}: p={p.tolist()} msm={msm} qlen={len(q)} tlen={len(t)} w={w} eb={eb} zd={zd} flag={flag:#x}"
        if msm and len(q) * len(t) > msm:
            branches["skip"] += 1
        elif cells == 0:
            branches["fast"] += 1
        else:
            branches["dp"] += 1
    assert min(branches.values()) > 100, branches


def test_fast_paths_by_definition():
    """An exact extension takes the ungapped path: score = max = a * qlen, one M of qlen, reach_end;
    an equal-length global with one substitution takes it too, with score (qlen-1)a - b."""
    p = O.align_params()
    mat = O.gen_nt4_mat(2, 8, 0, 1)
    rng = np.random.default_rng(1)
    for qlen in (1, 37, 150):
        q = rng.integers(0, 4, qlen).astype(np.uint8)
        t = np.concatenate([q, rng.integers(0, 4, 40).astype(np.uint8)])
        r = O.orc_align_pair(p, 0, q, t, mat, 101, 5, 100, 0x40)
        assert r["score"] == r["max"] == 2 * qlen and r["max_q"] == r["max_t"] == qlen - 1
        assert r["reach_end"] == 1 and r["mqe"] == 2 * qlen and r["mqe_t"] == qlen - 1 and r["cigar"] == [qlen << 4]
        t2 = q.copy()
        t2[qlen // 2] = (t2[qlen // 2] + 1) % 4
        g = O.orc_align_pair(p, 0, q, t2, mat, 101, -1, 100, 0)
        assert g["score"] == 2 * (qlen - 1) - 8 and g["max"] == 0 and g["cigar"] == [qlen << 4]
