"""GPU parity of mbg_ll_batch against the oracle's ksw_ll (itself pinned against ksw_ll_u8_core /
ksw_ll_i16_core and textbook local Smith-Waterman by the CPU tests).

Jobs are fuzzed over query lengths that fill every register width of the kernel (1..16 positions
a lane, up to the 1024-base limit), target lengths from 0, N bases, the matrices mate rescue builds
(pe.c:258: match 1, mismatch, transitions, methylation variants) and others, gap pairs with
gapo >= 1, SUBO thresholds (including 0: every row enters the b array), STOP scores and u8
saturation.  All five ksw_llrst_t fields must be equal.
"""
import numpy as np
import pytest

from minibwa_amd import mbg

O = pytest.importorskip("oracle.oracle")

pytestmark = pytest.mark.gpu

FIELDS = ("score", "te", "qe", "score2", "te2")


def _case(rng, qmax, tmax):
    ql = int(rng.integers(1, qmax))
    tl = int(rng.integers(0, tmax))
    q = rng.integers(0, 4, ql).astype(np.uint8)
    if rng.random() < 0.3:
        q[rng.random(ql) < 0.03] = 4
    t = rng.integers(0, 4, tl).astype(np.uint8)
    if tl and rng.random() < 0.8:
        for _ in range(1 + (rng.random() < 0.3)):
            st = int(rng.integers(0, max(1, tl - ql // 2)))
            cp = q.copy()
            mut = rng.random(ql) < rng.choice([0.0, 0.02, 0.1])
            cp[mut] = rng.integers(0, 4, int(mut.sum()))
            n = min(ql, tl - st)
            t[st:st + n] = cp[:n]
    if rng.random() < 0.2:
        t[rng.random(tl) < 0.02] = 4
    return q, t


def _check(ctx, Q, T, sizes, mat, gapo, gape, xtra):
    res = ctx.ksw_ll(Q, T, sizes, mat, gapo, gape, xtra)
    for i in range(len(Q)):
        sz = int(sizes[i]) if np.ndim(sizes) else int(sizes)
        want = O.orc_ll(Q[i], T[i], mat, sz, gapo, gape, xtra)
        got = {f: int(res[f][i]) for f in FIELDS}
        assert got == want, (i, len(Q[i]), len(T[i]), sz, gapo, gape, hex(xtra), got, want)
    return res


SETS = [  # (mat, gapo, gape, xtra)
    (O.gen_nt4_mat(1, 4, 0, 1), 6, 1, mbg.KSW_LL_SUBO | 19),          # mate rescue defaults (-A2 -B8 -O12 -E2)
    (O.gen_nt4_mat(1, 4, 0, 1), 6, 1, mbg.KSW_LL_SUBO | 0),
    (O.gen_nt4_mat(1, 2, 1, 1, 1), 3, 1, mbg.KSW_LL_SUBO | 10),       # transitions + C->T methylation
    (O.gen_nt4_mat(1, 3, 0, 1, 2), 1, 0, 0),
    (O.gen_nt4_mat(2, 5, 0, 2), 10, 3, mbg.KSW_LL_STOP | 90),
    (O.gen_nt4_mat(3, 2, 0, 1), 2, 2, mbg.KSW_LL_SUBO | mbg.KSW_LL_STOP | 40),
]


@pytest.mark.parametrize("size", [1, 2])
@pytest.mark.parametrize("set_i", range(len(SETS)))
def test_ll_fuzz(gpu_ctx, size, set_i):
    mat, gapo, gape, xtra = SETS[set_i]
    rng = np.random.default_rng(100 * size + set_i)
    Q, T = [], []
    for it in range(400):
        qmax = [40, 140, 300, 1025][it % 4] if size == 2 else [40, 140, 300, 700][it % 4]
        q, t = _case(rng, qmax, 900)
        if size == 2 and int(mat.max()) * min(len(q), len(t)) > 32767:
            continue
        Q.append(q)
        T.append(t)
    res = _check(gpu_ctx, Q, T, size, mat, gapo, gape, xtra)
    if xtra & mbg.KSW_LL_SUBO and not xtra & mbg.KSW_LL_STOP:
        assert (res["score2"] >= 0).sum() > 20
    if size == 1 and set_i == 0:
        assert (res["score"] == 255).sum() > 10


def test_ll_mixed_sizes_and_edges(gpu_ctx):
    """Both cores in one batch, empty targets and queries, one-base jobs, the 1024-base limit."""
    rng = np.random.default_rng(7)
    mat = O.gen_nt4_mat(1, 4, 0, 1)
    Q, T, S = [], [], []
    for ql, tl in [(0, 10), (0, 0), (1, 0), (1, 1), (5, 0), (1024, 300), (1024, 1), (16, 16), (17, 3), (64, 64),
                   (65, 90), (150, 550), (151, 300), (256, 256), (257, 260)]:
        for sz in (1, 2):
            q = rng.integers(0, 4, ql).astype(np.uint8)
            t = np.concatenate([rng.integers(0, 4, tl // 3), q[:tl - tl // 3]]).astype(np.uint8)[:tl]
            Q.append(q), T.append(t), S.append(sz)
    for it in range(300):
        q, t = _case(rng, 400, 600)
        Q.append(q), T.append(t), S.append(int(rng.integers(1, 3)))
    res = gpu_ctx.ksw_ll(Q, T, np.array(S), mat, 6, 1, mbg.KSW_LL_SUBO | 19)
    for i in range(len(Q)):
        if len(Q[i]) == 0:   # the reference reads before its buffers for an empty query
            assert tuple(int(res[f][i]) for f in FIELDS) == (0, -1, -1, -1, -1)
            continue
        want = O.orc_ll(Q[i], T[i], mat, S[i], 6, 1, mbg.KSW_LL_SUBO | 19)
        assert {f: int(res[f][i]) for f in FIELDS} == want, (i, len(Q[i]), len(T[i]), S[i])


def test_ll_many_jobs(gpu_ctx):
    """A mate-rescue-shaped batch: 20k reads of 150 bases against windows of 400-700 bases."""
    rng = np.random.default_rng(11)
    mat = O.gen_nt4_mat(1, 4, 0, 1)
    Q, T = [], []
    for it in range(20000):
        q = rng.integers(0, 4, 150).astype(np.uint8)
        tl = int(rng.integers(400, 700))
        t = rng.integers(0, 4, tl).astype(np.uint8)
        if it % 3:
            st = int(rng.integers(0, tl - 150))
            cp = q.copy()
            mut = rng.random(150) < 0.02
            cp[mut] = (cp[mut] + 1) % 4
            t[st:st + 150] = cp
        Q.append(q), T.append(t)
    res = gpu_ctx.ksw_ll(Q, T, 1, mat, 6, 1, mbg.KSW_LL_SUBO | 19)
    for i in rng.choice(len(Q), 1500, replace=False):
        want = O.orc_ll(Q[i], T[i], mat, 1, 6, 1, mbg.KSW_LL_SUBO | 19)
        assert {f: int(res[f][i]) for f in FIELDS} == want, i


def test_ll_unsupported(gpu_ctx):
    mat = O.gen_nt4_mat(1, 4, 0, 1)
    q = [np.zeros(10, np.uint8)]
    t = [np.zeros(20, np.uint8)]
    for args, what in [((q, t, 1, mat, 0, 1, 0), "gapo 0"),
                       (([np.zeros(1025, np.uint8)], t, 2, mat, 6, 1, 0), "qlen 1025"),
                       (([np.zeros(300, np.uint8)], [np.zeros(300, np.uint8)], 2, O.gen_nt4_mat(127, 4), 6, 1, 0),
                        "i16 saturation"),
                       ((q, t, 1, np.abs(mat), 6, 1, 0), "u8 without a negative score")]:
        with pytest.raises(mbg.MbgError, match="unsupported|outside"):
            gpu_ctx.ksw_ll(*args)
            pytest.fail(what)
